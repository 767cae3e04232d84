/* CPU oracle for the xFlow projection-step hot path — TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library, and only as the checker / the timed CPU baseline.  The product
 * library (libxflow_hip.so) neither links nor calls it.
 *
 * Each function restates one reference routine; the routine is cited next to
 * it.  Arithmetic is written in the reference's order (same operands, same
 * association) and the library is compiled with -ffp-contract=off, so for the
 * same inputs the results are bit-identical to the compiled reference.
 * Pinning (what checks each function against something independent):
 *   xo_space_mask, xo_rn_number, xo_activity, xo_apply(_dot), xo_dot,
 *   xo_absmax, xo_cg3, xo_bicg, xo_mr, xo_bicgstab, xo_thomas, xo_lod,
 *   xo_ib_interp, xo_ib_spread, xo_ib_index:
 *       bit-exact against the compiled reference's golden vectors
 *       (tests/golden), plus closed forms in tests/test_oracle_math.py;
 *   xo_cg_hs: textbook Hestenes-Stiefel CG (not in the reference), pinned by
 *       a dense numpy solve on small systems and by agreement with xo_cg3.
 *   xo_pack / xo_pack_apply: T1dPackOperator's packed rows and product,
 *       checked bit-exactly against the compiled reference's rows;
 *   xo_pack_solve: the reference solvers on those rows, bit-exact against
 *       the compiled reference's solutions on the packed operator.
 *   xo_cg_cg: single-reduction CG after Chronopoulos and Gear (J. Comput.
 *       Appl. Math. 25, 1989; not in the reference) with s = A p recomputed,
 *       the throughput solver's recurrence; pinned by a dense numpy solve and
 *       by agreement with xo_cg_hs (the same method in exact arithmetic).
 *   xo_transport_velocity, xo_pressure_rhs, xo_velocity_correct,
 *   xo_divergence_max: the code around the pressure solve inside
 *       TVPProblem::SolveBySplitting, which cannot be driven on its own; pinned
 *       by closed forms in tests/test_oracle_math.py (linear fields give exact
 *       face averages, divergences and gradients; a constant density of 2
 *       gives the rho^2 factor of XProblem.cpp:5495; Defined and PreDefined
 *       velocity points take the boundary value and the mirror).
 *   xo_dot, xo_absmax, xo_apply_dot: bit-exact against the golden vectors of
 *       the masked vector kernels (XSpace.cpp:2293-3487); xo_lincomb, xo_copy:
 *       inside every solver whose iterates are bit-exact against golden
 *       vectors, and their masking and zero-coefficient cases by closed form.
 *
 * Dot products are serial sums in i->j->k order, as in the reference
 * (XSpace.cpp:3219-3342).  Element-wise loops may run on several OpenMP
 * threads: that does not change any result bit.
 */
#define _GNU_SOURCE
#include "xf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define IDX(g, i, j, k) ((((int64_t)(i)) * (g)->l + (j)) * (g)->m + (k))

void xo_set_threads(int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
}

/* ---------------------------------------------------------------- masks */

/* Space mask from point types: XProblem.cpp:4155-4182 (same rule at 3722-3748). */
int xo_space_mask(int64_t ncell, const int32_t *codes, int32_t *sm) {
  for (int64_t c = 0; c < ncell; c++) {
    int v = codes[c];
    if (v == 0 || v == 1) sm[c] = v;
    else if (v == 3 || v == 7) sm[c] = 0;
    else if (v == 4 || v == 5 || v == 6 || v == 8) sm[c] = 1;
    else return XO_EINVAL;
  }
  return XO_OK;
}

/* Rn numbering: 1-based over mask>0 points in i->j->k order, 0 elsewhere;
 * masks must be 0/1.  XSpaceModel.cpp:276-411 (check 310-313, numbering 381-406). */
int64_t xo_rn_number(int n, int l, int m, const int32_t *sm, int32_t *rn) {
  int64_t ncell = (int64_t)n * l * m, num = 0;
  for (int64_t c = 0; c < ncell; c++)
    if (sm[c] != 0 && sm[c] != 1) return XO_EINVAL;
  for (int64_t c = 0; c < ncell; c++) rn[c] = sm[c] > 0 ? (int32_t)(++num) : 0;
  return num;
}

/* 7-point activity table: row rn-1 = [count, rn, neighbours i-1,i+1,j-1,j+1,
 * k-1,k+1 that are in bounds with mask>0]; unused slots stay 0.
 * XOperator.cpp:837-936. */
int xo_activity(int n, int l, int m, const int32_t *sm, const int32_t *rn, int32_t *act) {
  static const int d[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
  for (int i = 0; i < n; i++)
    for (int j = 0; j < l; j++)
      for (int k = 0; k < m; k++) {
        int64_t c = ((int64_t)i * l + j) * m + k;
        if (sm[c] <= 0) continue;
        int32_t *row = act + (int64_t)(rn[c] - 1) * 8;
        memset(row, 0, 8 * sizeof(int32_t));
        int count = 1;
        row[1] = rn[c];
        for (int q = 0; q < 6; q++) {
          int a = i + d[q][0], b = j + d[q][1], e = k + d[q][2];
          if (a >= 0 && a < n && b >= 0 && b < l && e >= 0 && e < m) {
            int64_t cc = ((int64_t)a * l + b) * m + e;
            if (sm[cc] > 0) row[++count] = rn[cc];
          }
        }
        row[0] = count;
      }
  return XO_OK;
}

/* ------------------------------------------------------------- operator */

/* One row of T3dLaplaceOperator::operator() (XOperator.cpp:938-1129) or, when
 * g->rho is set, of T3dVariableDivergentedOperator::operator()
 * (XOperator.cpp:1901-2233).  Returns XO_EINVAL where the reference throws. */
static int row_value(const xo_grid *g, const double *f, int i, int j, int k, double *out) {
  const int32_t *mk = g->codes;
  const double *hx = g->hx, *hy = g->hy, *hz = g->hz;
  int64_t c = IDX(g, i, j, k);
  int code = mk[c];
  int64_t sI = (int64_t)g->l * g->m, sJ = g->m;
  if (code == 1) {
    if (i < 1 || j < 1 || k < 1 || i > g->n - 2 || j > g->l - 2 || k > g->m - 2) return XO_EINVAL;
    double hx_2X = (hx[i - 1] + hx[i]) / 2;
    double hx_2Y = (hy[j - 1] + hy[j]) / 2;
    double hx_2Z = (hz[k - 1] + hz[k]) / 2;
    if (!g->rho) {
      double f_i_1 = (f[c] - f[c - sI]) / (hx[i - 1]);
      double f_i = (f[c + sI] - f[c]) / (hx[i]);
      double f_j_1 = (f[c] - f[c - sJ]) / (hy[j - 1]);
      double f_j = (f[c + sJ] - f[c]) / (hy[j]);
      double f_k_1 = (f[c] - f[c - 1]) / (hz[k - 1]);
      double f_k = (f[c + 1] - f[c]) / (hz[k]);
      if (mk[c - sI] == 0) f_i_1 = 0;
      if (mk[c + sI] == 0) f_i = 0;
      if (mk[c - sJ] == 0) f_j_1 = 0;
      if (mk[c + sJ] == 0) f_j = 0;
      if (mk[c - 1] == 0) f_k_1 = 0;
      if (mk[c + 1] == 0) f_k = 0;
      *out = -((f_i - f_i_1) / hx_2X + (f_j - f_j_1) / hx_2Y + (f_k - f_k_1) / hx_2Z);
    } else {
      const double *rho = g->rho;
      double f_i_1 = rho[c - sI] * (f[c] - f[c - sI]) / (hx[i - 1]);
      double f_i = rho[c + sI] * (f[c + sI] - f[c]) / (hx[i]);
      double f_j_1 = rho[c - sJ] * (f[c] - f[c - sJ]) / (hy[j - 1]);
      double f_j = rho[c + sJ] * (f[c + sJ] - f[c]) / (hy[j]);
      double f_k_1 = rho[c - 1] * (f[c] - f[c - 1]) / (hz[k - 1]);
      double f_k = rho[c + 1] * (f[c + 1] - f[c]) / (hz[k]);
      /* the density differences divide by the lower step (XOperator.cpp:1975-1977) */
      double d_i = (rho[c + sI] - rho[c]) / (hx[i - 1]);
      double d_j = (rho[c + sJ] - rho[c]) / (hy[j - 1]);
      double d_k = (rho[c + 1] - rho[c]) / (hz[k - 1]);
      if (mk[c - sI] == 0) f_i_1 = 0;
      if (mk[c + sI] == 0) { f_i = 0; d_i = 0; }
      if (mk[c - sJ] == 0) f_j_1 = 0;
      if (mk[c + sJ] == 0) { f_j = 0; d_j = 0; }
      if (mk[c - 1] == 0) f_k_1 = 0;
      if (mk[c + 1] == 0) { f_k = 0; d_k = 0; }
      double lap = (f_i - f_i_1) / hx_2X + (f_j - f_j_1) / hx_2Y + (f_k - f_k_1) / hx_2Z;
      double divd = d_i + d_j + d_k;
      /* a comment inside the reference's return expression leaves
       * -(rho*lap - divd*f) (XOperator.cpp:2025) */
      *out = -(rho[c] * lap - divd * f[c]);
    }
    return XO_OK;
  }
  if (code == 4 || code == 6) {
    if (!g->normals || !g->has_normal || !g->has_normal[c]) return XO_EINVAL;
    double nx = g->normals[3 * c], ny = g->normals[3 * c + 1], nz = g->normals[3 * c + 2];
    if (nx == 0 && ny == 0 && nz == 0) return XO_EINVAL;
    if (code == 4) { /* XOperator.cpp:1034-1069, 2027-2112 */
      if (nx != 0) { if (ny != 0 || nz != 0 || fabs(nx) != 1) return XO_EINVAL; }
      else if (ny != 0) { if (nx != 0 || nz != 0 || fabs(ny) != 1) return XO_EINVAL; }
      else if (nz != 0) { if (nx != 0 || ny != 0 || fabs(nz) != 1) return XO_EINVAL; }
      double s = 0;
      if (nx > 0) s = (f[c] + f[c - sI]) / 2;
      else if (nx < 0) s = (f[c + sI] + f[c]) / 2;
      if (ny > 0) s = (f[c] + f[c - sJ]) / 2;
      else if (ny < 0) s = (f[c + sJ] + f[c]) / 2;
      if (nz > 0) s = (f[c] + f[c - 1]) / 2;
      else if (nz < 0) s = (f[c + 1] + f[c]) / 2;
      *out = s;
      return XO_OK;
    }
    /* PreNormal: one-sided normal derivative of 4-point face averages,
     * XOperator.cpp:1070-1119 (Laplace) and 2113-2224 (divergented). */
    int sI_ = 0, sI_1 = 0, sJ_ = 0, sJ_1 = 0, sK_ = 0, sK_1 = 0;
    double hxl = 1, hyl = 1, hzl = 1;
    if (nx > 0) { sI_ = 0; sI_1 = -1; hxl = hx[i - 1]; }
    else if (nx < 0) { sI_ = 1; sI_1 = 0; hxl = hx[i]; }
    if (ny > 0) { sJ_ = 0; sJ_1 = -1; hyl = hy[j - 1]; }
    else if (ny < 0) { sJ_ = 1; sJ_1 = 0; hyl = hy[j]; }
    if (nz > 0) { sK_ = 0; sK_1 = -1; hzl = hz[k - 1]; }
    else if (nz < 0) { sK_ = 1; sK_1 = 0; hzl = hz[k]; }
#define F(a, b, e) f[IDX(g, i + (a), j + (b), k + (e))]
    double fl = (F(sI_, 0, 0) + F(sI_, sJ_, 0) + F(sI_, 0, sK_) + F(sI_, sJ_, sK_)) / 4;
    double fl_1 = (F(sI_1, 0, 0) + F(sI_1, sJ_, 0) + F(sI_1, 0, sK_) + F(sI_1, sJ_, sK_)) / 4;
    double dfx = (fl - fl_1) / hxl;
    fl = (F(0, sJ_, 0) + F(sI_, sJ_, 0) + F(0, sJ_, sK_) + F(sI_, sJ_, sK_)) / 4;
    fl_1 = (F(0, sJ_1, 0) + F(sI_, sJ_1, 0) + F(0, sJ_1, sK_) + F(sI_, sJ_1, sK_)) / 4;
    double dfy = (fl - fl_1) / hyl;
    fl = (F(0, 0, sK_) + F(sI_, 0, sK_) + F(0, sJ_, sK_) + F(sI_, sJ_, sK_)) / 4;
    fl_1 = (F(0, 0, sK_1) + F(sI_, 0, sK_1) + F(0, sJ_, sK_1) + F(sI_, sJ_, sK_1)) / 4;
    double dfz = (fl - fl_1) / hzl;
#undef F
    if (!g->rho) {
      *out = dfx * nx + dfy * ny + dfz * nz;
    } else {
      /* The divergented PreNormal row reads steps and densities at i-1 and
       * i+1 unconditionally, so it is only defined away from the array
       * faces; it also keeps the reference's precedence,
       * rho[i+1] - rho[i]/h (XOperator.cpp:2211-2213). */
      if (i < 1 || j < 1 || k < 1 || i > g->n - 2 || j > g->l - 2 || k > g->m - 2) return XO_EINVAL;
      const double *rho = g->rho;
      double d_x = rho[c + sI] - rho[c] / (hx[i - 1]);
      double d_y = rho[c + sJ] - rho[c] / (hy[j - 1]);
      double d_z = rho[c + 1] - rho[c] / (hz[k - 1]);
      double lap = dfx * nx + dfy * ny + dfz * nz;
      double divd = d_x + d_y + d_z;
      *out = rho[c] * lap - divd * f[c];
    }
    return XO_OK;
  }
  return XO_EINVAL;
}

/* y = A x (sign_b = 0), A x + b (+1) or A x - b (-1) at unknowns; other
 * entries of y are left as they are.  XSpace.cpp:3013-3217. */
int xo_apply(const xo_grid *g, const double *x, const double *b, double sign_b, double *y) {
  int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
  for (int i = 0; i < g->n; i++)
    for (int j = 0; j < g->l; j++)
      for (int k = 0; k < g->m; k++) {
        int64_t c = IDX(g, i, j, k);
        if (g->space_mask[c] <= 0) continue;
        double v;
        if (row_value(g, x, i, j, k, &v) != XO_OK) { bad = 1; continue; }
        if (sign_b > 0) y[c] = v + b[c];
        else if (sign_b < 0) y[c] = v - b[c];
        else y[c] = v;
      }
  return bad ? XO_EINVAL : XO_OK;
}

/* (A x, x): XSpace.cpp:3344-3419, serial i->j->k sum. */
double xo_apply_dot(const xo_grid *g, const double *x, int *status) {
  double res = 0;
  int st = XO_OK;
  for (int i = 0; i < g->n; i++)
    for (int j = 0; j < g->l; j++)
      for (int k = 0; k < g->m; k++) {
        int64_t c = IDX(g, i, j, k);
        if (g->space_mask[c] <= 0) continue;
        double v;
        if (row_value(g, x, i, j, k, &v) != XO_OK) { st = XO_EINVAL; continue; }
        res = res + v * x[c];
      }
  if (status) *status = st;
  return res;
}

/* res = p1*s1 + p2*s2 at unknowns (XSpace.cpp:2489-3010).  The reference's
 * special cases (p = 0, 1, -1) give the same values as the general formula. */
void xo_lincomb(const xo_grid *g, double p1, const double *s1, double p2, const double *s2, double *res) {
  int64_t nc = (int64_t)g->n * g->l * g->m;
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < nc; c++) {
    if (g->space_mask[c] <= 0) continue;
    if (p1 != 0 && p2 != 0) res[c] = p1 * s1[c] + p2 * s2[c];
    else if (p1 != 0) res[c] = p1 * s1[c];
    else if (p2 != 0) res[c] = p2 * s2[c];
    else res[c] = 0;
  }
}

/* res |= s at unknowns (XSpace.cpp:2293-2363, isCopying == false). */
void xo_copy(const xo_grid *g, const double *s, double *res) {
  int64_t nc = (int64_t)g->n * g->l * g->m;
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < nc; c++)
    if (g->space_mask[c] > 0) res[c] = s[c];
}

/* (a, b): XSpace.cpp:3261-3342, serial i->j->k sum of a*b. */
double xo_dot(const xo_grid *g, const double *a, const double *b) {
  int64_t nc = (int64_t)g->n * g->l * g->m;
  double res = 0;
  for (int64_t c = 0; c < nc; c++)
    if (g->space_mask[c] > 0) res = res + a[c] * b[c];
  return res;
}

/* max |x| over unknowns with the Rn index of the first strict maximum
 * (index 1 if every entry is 0).  XSpace.cpp:3428-3487, XSpace.cpp:961-973. */
double xo_absmax(const xo_grid *g, const double *x, const int32_t *rn, int32_t *rn_index) {
  int64_t nc = (int64_t)g->n * g->l * g->m;
  double res = 0;
  int32_t idx = 1;
  for (int64_t c = 0; c < nc; c++) {
    if (g->space_mask[c] <= 0) continue;
    double v = fabs(x[c]);
    if (v > res) { res = v; idx = rn[c]; }
  }
  if (rn_index) *rn_index = idx;
  return res;
}

/* ------------------------------------------------------------ solvers */

static double *dup(const double *s, int64_t nc) {
  double *p = (double *)malloc(sizeof(double) * nc);
  memcpy(p, s, sizeof(double) * nc);
  return p;
}
static double *zeros(int64_t nc) { return (double *)calloc(nc, sizeof(double)); }
static double rsqrt_ref(double v) { return (double)sqrtl((long double)v); }

/* The operator a reference solver runs on: the grid operator over the
 * unknown cells of an n*l*m array (A^T taken as A: the grid operator has no
 * transpose, XOperator.h:302), or packed rows of A and A^T over Rn vectors,
 * every entry an unknown (T1dPackOperator, XOperator.h:322-370). */
typedef struct {
  int64_t nc;                 /* vector length                               */
  const int32_t *unk;         /* unknown flags per entry; NULL: every entry  */
  const xo_grid *g;           /* grid operator, or NULL for packed rows      */
  const int64_t *ptr[2];      /* packed rows of A ([0]) and A^T ([1])        */
  const double *vals[2];
  const int32_t *cols[2];
} xo_op;

#define UNK(o, c) (!(o)->unk || (o)->unk[c] > 0)

/* Packed row r summed in stored order (XOperator.cpp:1551-1568). */
static double op_row(const xo_op *o, int t, int64_t r, const double *x) {
  double res = 0;
  for (int64_t q = o->ptr[t][r]; q < o->ptr[t][r + 1]; q++) res = res + o->vals[t][q] * x[o->cols[t][q]];
  return res;
}

static int op_apply(const xo_op *o, int t, const double *x, const double *b, double sign_b, double *y) {
  if (o->g) return xo_apply(o->g, x, b, sign_b, y);
  for (int64_t r = 0; r < o->nc; r++) {
    double v = op_row(o, t, r, x);
    if (sign_b > 0) y[r] = v + b[r];
    else if (sign_b < 0) y[r] = v - b[r];
    else y[r] = v;
  }
  return XO_OK;
}

static double op_apply_dot(const xo_op *o, const double *x, int *status) {
  if (o->g) return xo_apply_dot(o->g, x, status);
  double res = 0;
  for (int64_t r = 0; r < o->nc; r++) res = res + op_row(o, 0, r, x) * x[r];
  if (status) *status = XO_OK;
  return res;
}

static void op_lincomb(const xo_op *o, double p1, const double *s1, double p2, const double *s2, double *res) {
  if (o->g) { xo_lincomb(o->g, p1, s1, p2, s2, res); return; }
  for (int64_t c = 0; c < o->nc; c++) {
    if (p1 != 0 && p2 != 0) res[c] = p1 * s1[c] + p2 * s2[c];
    else if (p1 != 0) res[c] = p1 * s1[c];
    else if (p2 != 0) res[c] = p2 * s2[c];
    else res[c] = 0;
  }
}

static void op_copy(const xo_op *o, const double *s, double *res) {
  if (o->g) { xo_copy(o->g, s, res); return; }
  memcpy(res, s, sizeof(double) * o->nc);
}

static double op_dot(const xo_op *o, const double *a, const double *b) {
  if (o->g) return xo_dot(o->g, a, b);
  double res = 0;
  for (int64_t c = 0; c < o->nc; c++) res = res + a[c] * b[c];
  return res;
}

static xo_op grid_op(const xo_grid *g) {
  xo_op o;
  memset(&o, 0, sizeof(o));
  o.nc = (int64_t)g->n * g->l * g->m;
  o.unk = g->space_mask;
  o.g = g;
  return o;
}

/* Samarskii three-term CG, XSLAE.cpp:56-157.  Stops when ||r||_2 <= eps
 * (absolute; the reference hard-codes 1e-5 at XSLAE.cpp:64).  maxit bounds
 * the loop (the reference has no bound); NaN counts as failure here. */
static int cg3_op(const xo_op *o, const double *b, const double *x0, double *x, double eps, int maxit,
                  int *iters, double *resid) {
  int64_t nc = o->nc;
  int st = XO_OK, it = 0;
  double *res = dup(x0, nc);
  double *xk_1 = dup(x0, nc), *rk_1 = zeros(nc), *xk = NULL, *rk = NULL;
  double rr1 = 0, eps_cur = 0;
  op_lincomb(o, 0, NULL, 0, NULL, res);
  if (op_apply(o, 0, xk_1, b, -1, rk_1)) { st = XO_EINVAL; goto done; }
  rr1 = op_dot(o, rk_1, rk_1);
  op_copy(o, xk_1, res);
  eps_cur = rsqrt_ref(rr1);
  if (rr1 != 0) {
    int s2;
    double tauk = rr1 / op_apply_dot(o, rk_1, &s2);
    xk = dup(x0, nc);
    op_lincomb(o, 1, xk_1, -tauk, rk_1, xk);
    rk = zeros(nc);
    op_apply(o, 0, xk, b, -1, rk);
    double rr = op_dot(o, rk, rk);
    op_copy(o, xk, res);
    eps_cur = rsqrt_ref(rr);
    if (rr != 0) {
      double alfak = 1;
      while (eps_cur > eps) {
        if (it >= maxit) { st = XO_EMAXIT; break; }
        it++;
        double tauk1 = rr / op_apply_dot(o, rk, &s2);
        double alfak1 = 1 - (1 / alfak) * (tauk1 / tauk) * (rr / rr1);
        alfak1 = 1 / alfak1;
        op_lincomb(o, alfak1, xk, 1 - alfak1, xk_1, res);
        op_lincomb(o, 1, res, -(tauk1 * alfak1), rk, res);
        op_copy(o, xk, xk_1);
        op_copy(o, rk, rk_1);
        rr1 = rr;
        tauk = tauk1;
        alfak = alfak1;
        op_copy(o, res, xk);
        op_apply(o, 0, xk, b, -1, rk);
        rr = op_dot(o, rk, rk);
        eps_cur = rsqrt_ref(rr);
      }
      if (eps_cur != eps_cur) st = XO_ENAN;
    }
  }
done:
  memcpy(x, res, sizeof(double) * nc);
  if (iters) *iters = it;
  if (resid) *resid = eps_cur;
  free(res); free(xk_1); free(rk_1); free(xk); free(rk);
  return st;
}

/* BiCG, XSLAE.cpp:159-348.  x0 is used only for its non-unknown entries
 * (the reference zeroes the unknowns, XSLAE.cpp:171).  The shadow sequence
 * applies the operator's transpose (op_apply with t = 1): the packed rows of
 * A^T, or A itself on the grid. */
static int bicg_op(const xo_op *o, const double *b, const double *x0, double *x, double eps, int maxit,
                   int *iters, double *resid) {
  int64_t nc = o->nc;
  int st = XO_OK, it = 0;
  double *apk = zeros(nc);
  double *xk = dup(x0, nc), *xk_new = dup(x0, nc);
  double *rk = zeros(nc), *rt = zeros(nc), *rk_new = zeros(nc), *rt_new = zeros(nc);
  double *pk = zeros(nc), *pt = zeros(nc), *pk_new = zeros(nc), *pt_new = zeros(nc);
  double *res = dup(x0, nc);
  op_lincomb(o, 0, NULL, 0, NULL, res);
  op_lincomb(o, 0, NULL, 0, NULL, xk);
  op_apply(o, 0, xk, b, -1, rk);
  op_lincomb(o, -1, rk, 0, NULL, rt);
  op_copy(o, rt, rk);
  op_copy(o, rk, pk);
  op_copy(o, rt, pt);
  double rr = op_dot(o, rk, rk);
  op_copy(o, xk, res);
  double eps_cur = rsqrt_ref(rr);
  while (eps_cur > eps) {
    if (it >= maxit) { st = XO_EMAXIT; break; }
    it++;
    double a1 = op_dot(o, rk, rt);
    /* (A(pk), pt): fused product of A pk with pt, serial i->j->k sum */
    op_apply(o, 0, pk, NULL, 0, apk);
    double a2 = 0;
    for (int64_t c = 0; c < nc; c++)
      if (UNK(o, c)) a2 = a2 + apk[c] * pt[c];
    double alpha = a1 / a2;
    op_lincomb(o, 1, xk, alpha, pk, xk_new);
    op_apply(o, 0, pk, NULL, 0, rk_new);
    op_lincomb(o, -alpha, rk_new, 1, rk, rk_new);
    op_apply(o, 1, pt, NULL, 0, rt_new);
    op_lincomb(o, -alpha, rt_new, 1, rt, rt_new);
    double b1 = op_dot(o, rk_new, rt_new);
    double b2 = op_dot(o, rk, rt);
    if (b1 == 0 || b2 == 0) { st = XO_EBREAKDOWN; break; }
    double beta = b1 / b2;
    op_lincomb(o, 1, rk_new, beta, pk, pk_new);
    op_lincomb(o, 1, rt_new, beta, pt, pt_new);
    op_copy(o, xk_new, xk);
    op_copy(o, rk_new, rk);
    op_copy(o, rt_new, rt);
    op_copy(o, pk_new, pk);
    op_copy(o, pt_new, pt);
    rr = op_dot(o, rk, rk);
    eps_cur = rsqrt_ref(rr);
  }
  if (eps_cur != eps_cur) st = XO_ENAN;
  op_copy(o, xk, res);
  memcpy(x, res, sizeof(double) * nc);
  if (iters) *iters = it;
  if (resid) *resid = eps_cur;
  free(xk); free(xk_new); free(rk); free(rt); free(rk_new); free(rt_new);
  free(pk); free(pt); free(pk_new); free(pt_new); free(res); free(apk);
  return st;
}

/* Minimal residual iteration, XSLAE.cpp:350-414: alpha = -(Ar,r)/(Ar,Ar). */
static int mr_op(const xo_op *o, const double *b, const double *x0, double *x, double eps, int maxit,
                 int *iters, double *resid) {
  int64_t nc = o->nc;
  int st = XO_OK, it = 0;
  double *xk = dup(x0, nc), *rk = zeros(nc), *ark = zeros(nc), *res = dup(x0, nc);
  op_lincomb(o, 0, NULL, 0, NULL, res);
  op_apply(o, 0, xk, b, -1, rk);
  double rr = op_dot(o, rk, rk);
  op_copy(o, xk, res);
  double eps_cur = rsqrt_ref(rr);
  if (rr != 0) {
    while (eps_cur > eps) {
      if (it >= maxit) { st = XO_EMAXIT; break; }
      it++;
      op_apply(o, 0, rk, NULL, 0, ark);
      double alfa = (-1) * op_dot(o, ark, rk) / op_dot(o, ark, ark);
      op_lincomb(o, 1, xk, alfa, rk, xk);
      op_apply(o, 0, xk, b, -1, rk);
      rr = op_dot(o, rk, rk);
      eps_cur = rsqrt_ref(rr);
    }
    if (eps_cur != eps_cur) st = XO_ENAN;
  }
  op_copy(o, xk, res);
  memcpy(x, res, sizeof(double) * nc);
  if (iters) *iters = it;
  if (resid) *resid = eps_cur;
  free(xk); free(rk); free(ark); free(res);
  return st;
}

/* BiCGStab, XSLAE.cpp:667-767; stops on max|r| <= eps (L-infinity). */
static int bicgstab_op(const xo_op *o, const double *b, const double *x0, double *x, double eps, int maxit,
                       int *iters, double *resid) {
  int64_t nc = o->nc;
  int st = XO_OK, it = 0;
  double *Un = dup(x0, nc), *tmp1 = zeros(nc), *tmp2 = zeros(nc), *Rn = zeros(nc), *sRn = zeros(nc);
  double *Pn = zeros(nc), *Pn1 = zeros(nc), *Vn = zeros(nc), *Sn = zeros(nc), *Tn = zeros(nc);
  double *res = dup(x0, nc);
  op_lincomb(o, 0, NULL, 0, NULL, res);
  op_apply(o, 0, Un, NULL, 0, Rn);
  op_lincomb(o, 1, b, -1, Rn, Rn);
  op_copy(o, Rn, sRn);
  double pn = 1., an = 1., bn = 1., wn = 1.;
  double norm = 0;
  {
    int64_t c;
    for (c = 0; c < nc; c++) if (UNK(o, c) && fabs(Rn[c]) > norm) norm = fabs(Rn[c]);
  }
  while (norm > eps) {
    if (it >= maxit) { st = XO_EMAXIT; break; }
    it++;
    double pn1 = op_dot(o, sRn, Rn);
    bn = (pn1 / pn) * (an / wn);
    op_lincomb(o, wn, Vn, 0, NULL, Pn1);
    op_lincomb(o, 1, Pn, -1, Pn1, Pn1);
    op_lincomb(o, bn, Pn1, 0, NULL, Pn1);
    op_lincomb(o, 1, Rn, 1, Pn1, Pn1);
    op_apply(o, 0, Pn1, NULL, 0, Vn);
    an = pn1 / op_dot(o, sRn, Vn);
    op_lincomb(o, an, Vn, 0, NULL, Sn);
    op_lincomb(o, 1, Rn, -1, Sn, Sn);
    op_apply(o, 0, Sn, NULL, 0, Tn);
    wn = op_dot(o, Tn, Sn) / op_dot(o, Tn, Tn);
    op_lincomb(o, an, Pn1, 0, NULL, tmp1);
    op_lincomb(o, 1, tmp1, wn, Sn, tmp2);
    op_lincomb(o, 1, Un, 1, tmp2, Un);
    op_lincomb(o, wn, Tn, 0, NULL, Rn);
    op_lincomb(o, 1, Sn, -1, Rn, Rn);
    pn = pn1;
    op_copy(o, Pn1, Pn);
    norm = 0;
    for (int64_t c = 0; c < nc; c++) if (UNK(o, c) && fabs(Rn[c]) > norm) norm = fabs(Rn[c]);
  }
  if (norm != norm) st = XO_ENAN;
  op_copy(o, Un, res);
  memcpy(x, res, sizeof(double) * nc);
  if (iters) *iters = it;
  if (resid) *resid = norm;
  free(Un); free(tmp1); free(tmp2); free(Rn); free(sRn); free(Pn); free(Pn1); free(Vn); free(Sn);
  free(Tn); free(res);
  return st;
}

int xo_cg3(const xo_grid *g, const double *b, const double *x0, double *x, double eps, int maxit,
           int *iters, double *resid) {
  xo_op o = grid_op(g);
  return cg3_op(&o, b, x0, x, eps, maxit, iters, resid);
}

int xo_bicg(const xo_grid *g, const double *b, const double *x0, double *x, double eps, int maxit,
            int *iters, double *resid) {
  xo_op o = grid_op(g);
  return bicg_op(&o, b, x0, x, eps, maxit, iters, resid);
}

int xo_mr(const xo_grid *g, const double *b, const double *x0, double *x, double eps, int maxit,
          int *iters, double *resid) {
  xo_op o = grid_op(g);
  return mr_op(&o, b, x0, x, eps, maxit, iters, resid);
}

int xo_bicgstab(const xo_grid *g, const double *b, const double *x0, double *x, double eps, int maxit,
                int *iters, double *resid) {
  xo_op o = grid_op(g);
  return bicgstab_op(&o, b, x0, x, eps, maxit, iters, resid);
}

/* A reference solver (0 CG3, 1 BiCG, 2 minimal residuals, 3 BiCGStab) on
 * packed rows of A and A^T (xo_pack's layout) over Rn vectors of length nunk;
 * BiCG's shadow sequence takes the transposed rows, as the reference's copy
 * of the packed operator toggled by operator~ does (XSLAE.cpp:205-206). */
int xo_pack_solve(int method, int64_t nunk, const int32_t *counts, const double *vals, const int32_t *cols,
                  const int32_t *counts_t, const double *vals_t, const int32_t *cols_t, const double *b,
                  const double *x0, double *x, double eps, int maxit, int *iters, double *resid) {
  int64_t *ptr = (int64_t *)malloc(sizeof(int64_t) * 2 * (nunk + 1));
  const int32_t *cn[2] = {counts, counts_t};
  xo_op o;
  memset(&o, 0, sizeof(o));
  o.nc = nunk;
  for (int t = 0; t < 2; t++) {
    int64_t *pt = ptr + t * (nunk + 1);
    pt[0] = 0;
    for (int64_t r = 0; r < nunk; r++) pt[r + 1] = pt[r] + cn[t][r];
    o.ptr[t] = pt;
  }
  o.vals[0] = vals; o.vals[1] = vals_t;
  o.cols[0] = cols; o.cols[1] = cols_t;
  int st;
  switch (method) {
    case 0: st = cg3_op(&o, b, x0, x, eps, maxit, iters, resid); break;
    case 1: st = bicg_op(&o, b, x0, x, eps, maxit, iters, resid); break;
    case 2: st = mr_op(&o, b, x0, x, eps, maxit, iters, resid); break;
    case 3: st = bicgstab_op(&o, b, x0, x, eps, maxit, iters, resid); break;
    default: st = XO_EINVAL;
  }
  free(ptr);
  return st;
}

/* Textbook Hestenes-Stiefel CG on A e = b - A x0 (not a reference routine:
 * the GPU fast path's algorithm, used as its checker).  Stops when
 * ||r||_2 <= rtol * ||b - A x0||_2.  project_mean removes the mean over the
 * unknowns from the right-hand side and the iterates (pure-Neumann systems). */
int xo_cg_hs(const xo_grid *g, const double *b, const double *x0, double *x, double rtol, int maxit,
             int project_mean, int *iters, double *resid) {
  int64_t nc = (int64_t)g->n * g->l * g->m;
  double *r = zeros(nc), *p = zeros(nc), *q = zeros(nc), *e = zeros(nc);
  int st = XO_OK, it = 0;
  memcpy(x, x0, sizeof(double) * nc);
  xo_apply(g, x, b, -1, r);
  int64_t nu = 0;
  for (int64_t c = 0; c < nc; c++) {
    if (g->space_mask[c] > 0) { r[c] = -r[c]; nu++; } else r[c] = 0;
  }
  if (project_mean && nu) {
    double s = 0;
    for (int64_t c = 0; c < nc; c++) if (g->space_mask[c] > 0) s += r[c];
    s /= (double)nu;
    for (int64_t c = 0; c < nc; c++) if (g->space_mask[c] > 0) r[c] -= s;
  }
  double rr = xo_dot(g, r, r), r0 = sqrt(rr);
  memcpy(p, r, sizeof(double) * nc);
  while (sqrt(rr) > rtol * r0) {
    if (it >= maxit) { st = XO_EMAXIT; break; }
    it++;
    xo_apply(g, p, NULL, 0, q);
    double pq = xo_dot(g, p, q);
    double alpha = rr / pq;
    for (int64_t c = 0; c < nc; c++)
      if (g->space_mask[c] > 0) { e[c] += alpha * p[c]; r[c] -= alpha * q[c]; }
    double rr_new = xo_dot(g, r, r);
    double beta = rr_new / rr;
    rr = rr_new;
    for (int64_t c = 0; c < nc; c++)
      if (g->space_mask[c] > 0) p[c] = r[c] + beta * p[c];
    if (rr != rr) { st = XO_ENAN; break; }
  }
  if (project_mean && nu) {
    double s = 0;
    for (int64_t c = 0; c < nc; c++) if (g->space_mask[c] > 0) s += e[c];
    s /= (double)nu;
    for (int64_t c = 0; c < nc; c++) if (g->space_mask[c] > 0) e[c] -= s;
  }
  for (int64_t c = 0; c < nc; c++) if (g->space_mask[c] > 0) x[c] += e[c];
  if (iters) *iters = it;
  if (resid) *resid = r0 > 0 ? sqrt(rr) / r0 : 0;
  free(r); free(p); free(q); free(e);
  return st;
}

/* Single-reduction CG: per iteration w = A r, gamma = (r, r), delta = (w, r);
 * stop when gamma <= rtol^2 gamma_0; beta = gamma/gamma_old,
 * alpha = gamma/(delta - beta gamma/alpha_old) (alpha = gamma/delta first);
 * p = r + beta p, s = A p, e += alpha p, r -= alpha s.  (Chronopoulos and
 * Gear carry s = w + beta s instead; s = A p holds in exact arithmetic, and
 * the GPU solver recomputes A p rather than storing w and s.) */
int xo_cg_cg(const xo_grid *g, const double *b, const double *x0, double *x, double rtol, int maxit,
             int project_mean, int *iters, double *resid) {
  int64_t nc = (int64_t)g->n * g->l * g->m;
  double *r = zeros(nc), *p = zeros(nc), *s = zeros(nc), *w = zeros(nc), *e = zeros(nc);
  int st = XO_OK, it = 0;
  memcpy(x, x0, sizeof(double) * nc);
  xo_apply(g, x, b, -1, r);
  int64_t nu = 0;
  for (int64_t c = 0; c < nc; c++) {
    if (g->space_mask[c] > 0) { r[c] = -r[c]; nu++; } else r[c] = 0;
  }
  if (project_mean && nu) {
    double sm = 0;
    for (int64_t c = 0; c < nc; c++) if (g->space_mask[c] > 0) sm += r[c];
    sm /= (double)nu;
    for (int64_t c = 0; c < nc; c++) if (g->space_mask[c] > 0) r[c] -= sm;
  }
  const double rr0 = xo_dot(g, r, r), stop2 = rtol * rtol * rr0;
  double gam = rr0, gold = 0, aold = 0;
  if (!(rr0 <= stop2)) {
    for (;;) {
      xo_apply(g, r, NULL, 0, w);
      gam = xo_dot(g, r, r);
      double del = xo_dot(g, w, r);
      if (gam != gam || del != del) { st = XO_ENAN; break; }
      if (gam <= stop2) break;
      if (it >= maxit) { st = XO_EMAXIT; break; }
      double beta = 0, alpha;
      if (it == 0) alpha = gam / del;
      else {
        beta = gam / gold;
        alpha = gam / (del - beta * gam / aold);
      }
      for (int64_t c = 0; c < nc; c++)
        if (g->space_mask[c] > 0) p[c] = r[c] + beta * p[c];
      xo_apply(g, p, NULL, 0, s);
      for (int64_t c = 0; c < nc; c++)
        if (g->space_mask[c] > 0) {
          e[c] += alpha * p[c];
          r[c] -= alpha * s[c];
        }
      gold = gam;
      aold = alpha;
      it++;
    }
  }
  if (project_mean && nu) {
    double sm = 0;
    for (int64_t c = 0; c < nc; c++) if (g->space_mask[c] > 0) sm += e[c];
    sm /= (double)nu;
    for (int64_t c = 0; c < nc; c++) if (g->space_mask[c] > 0) e[c] -= sm;
  }
  for (int64_t c = 0; c < nc; c++) if (g->space_mask[c] > 0) x[c] += e[c];
  if (iters) *iters = it;
  if (resid) *resid = rr0 > 0 ? sqrt(gam / rr0) : 0;
  free(r); free(p); free(s); free(w); free(e);
  return st;
}

/* ------------------------------------------------------ packed operator */

/* T1dPackOperator (XOperator.cpp:1135-1345): the operator's matrix in packed
 * rows over the Rn unknowns.  Row r of A (transpose = 0) lists, in the order
 * of the activity-table row r (self, then the i-1, i+1, j-1, j+1, k-1, k+1
 * neighbours that are unknowns), the non-zero coefficients A[r][q]; row r of
 * A^T lists, in the same order over q, the non-zero A[q][r].  A coefficient is
 * the row evaluated on a unit vector, as the reference's column probing does
 * (XOperator.cpp:1207-1224).  counts[nunk], vals/cols sized 7*nunk; returns
 * the number of entries, or XO_EINVAL where a row throws or is empty. */
int64_t xo_pack(const xo_grid *g, int transpose, int32_t *counts, double *vals, int32_t *cols) {
  const int64_t nc = (int64_t)g->n * g->l * g->m;
  int32_t *rn = (int32_t *)malloc(sizeof(int32_t) * nc);
  int64_t nunk = xo_rn_number(g->n, g->l, g->m, g->space_mask, rn);
  int64_t *cell = (int64_t *)malloc(sizeof(int64_t) * (nunk > 0 ? nunk : 1));
  for (int64_t c = 0; c < nc; c++)
    if (rn[c] > 0) cell[rn[c] - 1] = c;
  double *f = zeros(nc);
  int64_t q = 0, st = 0;
  static const int d[7][3] = {{0, 0, 0}, {-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
  for (int64_t r = 0; r < nunk && st == 0; r++) {
    const int64_t c = cell[r];
    const int i = (int)(c / ((int64_t)g->l * g->m)), j = (int)((c / g->m) % g->l), k = (int)(c % g->m);
    int cnt = 0;
    for (int e = 0; e < 7; e++) {
      const int a = i + d[e][0], b = j + d[e][1], w = k + d[e][2];
      if (a < 0 || a >= g->n || b < 0 || b >= g->l || w < 0 || w >= g->m) continue;
      const int64_t cc = IDX(g, a, b, w);
      if (g->space_mask[cc] <= 0) continue;
      double v = 0;
      int s2;
      if (!transpose) {  /* A[r][rn(cc)]: row at c, unit vector at cc */
        f[cc] = 1;
        s2 = row_value(g, f, i, j, k, &v);
        f[cc] = 0;
      } else {           /* A[rn(cc)][r]: row at cc, unit vector at c */
        f[c] = 1;
        s2 = row_value(g, f, a, b, w, &v);
        f[c] = 0;
      }
      if (s2) { st = XO_EINVAL; break; }
      if (v != 0) {
        vals[q] = v;
        cols[q] = rn[cc] - 1;
        q++;
        cnt++;
      }
    }
    if (!cnt) st = XO_EINVAL;
    counts[r] = cnt;
  }
  free(rn); free(cell); free(f);
  return st ? st : q;
}

/* y[r] = sum over the packed row r in stored order (T1dPackOperator::operator(),
 * XOperator.cpp:1551-1568). */
void xo_pack_apply(int64_t nunk, const int32_t *counts, const double *vals, const int32_t *cols, const double *x,
                   double *y) {
  int64_t q = 0;
  for (int64_t r = 0; r < nunk; r++) {
    double res = 0;
    for (int e = 0; e < counts[r]; e++, q++) res = res + vals[q] * x[cols[q]];
    y[r] = res;
  }
}

/* --------------------------------------------------------- transport */

/* Thomas sweep with known ends u[0], u[M]: XSLAE.cpp:416-445. */
void xo_thomas(int M, double *u, const double *a, const double *b, const double *c, const double *f) {
  double *A = (double *)malloc(sizeof(double) * (M + 1));
  double *B = (double *)malloc(sizeof(double) * (M + 1));
  A[0] = 0;
  B[0] = u[0];
  A[M] = B[M] = 0;
  for (int q = 1; q < M; q++) {
    A[q] = (-c[q]) / (a[q] * A[q - 1] + b[q]);
    B[q] = (f[q] - a[q] * B[q - 1]) / (a[q] * A[q - 1] + b[q]);
  }
  for (int q = M - 1; q > 0; q--) u[q] = A[q] * u[q + 1] + B[q];
  free(A);
  free(B);
}

/* Explicit half-weight terms of one direction e (0=x,1=y,2=z) for the cell c,
 * built from field fd (stage or previous) and fp (previous):
 * XProblem.cpp:3246-3292 and its y/z counterparts. */
static void side_terms(const xo_grid *g, int e, int64_t c, int ii, const double *fd, const double *fp,
                       const double *vel, double kap, int scheme, double *FC, double *FD) {
  const double *h = e == 0 ? g->hx : e == 1 ? g->hy : g->hz;
  int64_t s = e == 0 ? (int64_t)g->l * g->m : e == 1 ? g->m : 1;
  double h_1 = h[ii - 1], hh = h[ii], hx2 = h[ii - 1] + h[ii];
  double hxhxh = h[ii - 1] * h[ii] * (h[ii - 1] + h[ii]) / 2;
  double fc = 0, fdd = 0;
  for (int item = 1; item <= 2; item++) {
    const double *f_ = item == 1 ? fd : fp;
    if (scheme == 0) fc = fc + 0.5 * (-vel[c]) * (f_[c + s] - f_[c - s]) / hx2;
    else {
      if (vel[c] < 0) fc = fc + 0.5 * (-vel[c]) * (f_[c + s] - f_[c]) / hh;
      else fc = fc + 0.5 * (-vel[c]) * (f_[c] - f_[c - s]) / hh;
    }
    fdd = fdd + 0.5 * (kap) * ((hh * f_[c - s] - hx2 * f_[c] + h_1 * f_[c + s]) / hxhxh);
  }
  *FC = fc;
  *FD = fdd;
}

/* LOD transport step THydrodynamicsProblem::SolveTransferEquation__,
 * XProblem.cpp:2784-3475 (direction order 1, all terms on).  f is updated in
 * place; f_prev is the previous time level.  Returns XO_EINVAL where the
 * reference throws (bad sector borders or sizes). */
int xo_lod(const xo_grid *g, double *f, const double *f_prev, const double *u, const double *v,
           const double *w, const double *rp, const double *kappa, double tau, int scheme) {
  int64_t nc = (int64_t)g->n * g->l * g->m;
  double *stage[3];
  stage[0] = dup(f_prev, nc);
  stage[1] = dup(f_prev, nc);
  stage[2] = f;
  const int32_t *mk = g->codes;
  int N = g->n - 1, L = g->l - 1, M = g->m - 1;
  const double *hs[3] = {g->hx, g->hy, g->hz};
  const double *vels[3] = {u, v, w};
  int64_t strides[3] = {(int64_t)g->l * g->m, g->m, 1};
  int ends[3] = {N, L, M};
  int maxline = N > L ? N : L;
  if (M > maxline) maxline = M;
  int cap = maxline + 4;
  double *sv = malloc(sizeof(double) * cap), *sa = malloc(sizeof(double) * cap);
  double *sb = malloc(sizeof(double) * cap), *scc = malloc(sizeof(double) * cap);
  double *srp = malloc(sizeof(double) * cap);
  int st = XO_OK;
  for (int dg = 1; dg < 4 && st == XO_OK; dg++) {
    for (int dfk = 1; dfk < 4 && st == XO_OK; dfk++) {
      int dir;
      if (dfk == 1) dir = dg;
      else if (dg == 1) dir = dfk;
      else if (dg == 2) dir = dfk == 2 ? 1 : 3;
      else dir = dfk - 1;
      int a = dir - 1;                    /* active axis */
      /* dir1 / dir2 axes as in XProblem.cpp:2871-2888 */
      int ax1 = a == 2 ? 1 : 2, ax2 = a == 0 ? 1 : 0;
      int64_t sa_ = strides[a];
      const double *h = hs[a];
      double *fa = stage[dg - 1];         /* stage written by this global direction */
      for (int d1 = 0; d1 <= ends[ax1] && st == XO_OK; d1++)
        for (int d2 = 0; d2 <= ends[ax2] && st == XO_OK; d2++) {
          int64_t base = (int64_t)d1 * strides[ax1] + (int64_t)d2 * strides[ax2];
          int beg = -1, mbeg = 0;
          for (int q = 0; q <= ends[a]; q++) {
            int mq = mk[base + q * sa_];
            if (beg < 0) {
              if (mq == 1) {
                if (q == 0) { st = XO_EINVAL; break; }
                beg = q - 1;
                mbeg = mk[base + beg * sa_];
              }
              continue;
            }
            if (mq == 1) continue;
            int end = q, mend = mq;
            if (mbeg == 1 || mbeg == 0 || mend == 1 || mend == 0) { st = XO_EINVAL; break; }
            if (end - beg < 2) { st = XO_EINVAL; break; }
            int64_t cb = base + beg * sa_, ce = base + end * sa_;
            if (dir != dg) {
              /* border-only recalculation (XProblem.cpp:2965-3033) */
              if (mbeg == 3 || mbeg == 7) fa[cb] = rp[cb];
              else if (mbeg == 4) fa[cb] = 2 * rp[cb] - fa[cb + sa_];
              else if (mbeg == 5 || mbeg == 6) fa[cb] = -(h[beg]) * rp[cb] + fa[cb + sa_];
              else { st = XO_EINVAL; break; }
              if (mend == 3 || mend == 7) fa[ce] = rp[ce];
              else if (mend == 4) fa[ce] = 2 * rp[ce] - fa[ce - sa_];
              else if (mend == 5 || mend == 6) fa[ce] = (h[end - 1]) * rp[ce] + fa[ce - sa_];
              else { st = XO_EINVAL; break; }
              beg = -1;
              continue;
            }
            int fl = (mbeg != 3 && mbeg != 7), fr = (mend != 3 && mend != 7);
            int dim = end - beg + 1 + fl + fr;
            int sd = dim - 1;
            sa[0] = 0; sa[sd] = 0; sb[0] = 0; sb[sd] = 0; scc[0] = 0; scc[sd] = 0; srp[0] = 0; srp[sd] = 0;
            sv[0] = 0; sv[sd] = 0;
            if (mbeg == 3 || mbeg == 7) sv[0] = rp[cb];
            else if (mbeg == 4) { sv[0] = 0; sa[1] = 1; sb[1] = 0.5; scc[1] = 0.5; srp[1] = rp[cb]; }
            else if (mbeg == 5 || mbeg == 6) {
              sv[0] = 0; sa[1] = 1; sb[1] = -1 / h[beg]; scc[1] = 1 / h[beg]; srp[1] = rp[cb];
            } else { st = XO_EINVAL; break; }
            if (mend == 3 || mend == 7) sv[sd] = rp[ce];
            else if (mend == 4) { sv[sd] = 0; scc[sd - 1] = 1; sa[sd - 1] = 0.5; sb[sd - 1] = 0.5; srp[sd - 1] = rp[ce]; }
            else if (mend == 5 || mend == 6) {
              sv[sd] = 0; scc[sd - 1] = 1; sa[sd - 1] = -1 / h[end - 1]; sb[sd - 1] = 1 / h[end - 1];
              srp[sd - 1] = rp[ce];
            } else { st = XO_EINVAL; break; }
            for (int s = 1 + fl; s < sd - fr; s++) {
              int da = beg + s - fl;
              int64_t c = base + (int64_t)da * sa_;
              if (mk[c] != 1) { st = XO_EINVAL; break; }
              int i = (int)(c / strides[0]), j = (int)((c / strides[1]) % g->l), k = (int)(c % g->m);
              int idx3[3] = {i, j, k};
              int ii = idx3[a];
              double h_1 = h[ii - 1], hh = h[ii], hx2 = h[ii - 1] + h[ii];
              double hxhxh = h[ii - 1] * h[ii] * (h[ii - 1] + h[ii]) / 2;
              double kap = kappa[c];
              const double *vt = vels[a];
              if (scheme == 0) {
                sa[s] = 1 * (0.5 * (-vt[c] / hx2)) + 1 * (0.5 * (-(kap)*hh / hxhxh));
                sb[s] = 1 / tau + 1 * (0.5 * (kap)*hx2 / hxhxh);
                scc[s] = 1 * (0.5 * (vt[c] / hx2)) + 1 * (0.5 * (-(kap)*h_1 / hxhxh));
              } else {
                if (vt[c] < 0) {
                  sa[s] = 1 * (0.5 * (-(kap)*hh / hxhxh));
                  sb[s] = 1 / tau + 1 * (0.5 * (-vt[c] / hh)) + 1 * (0.5 * (kap)*hx2 / hxhxh);
                  scc[s] = 1 * (0.5 * (vt[c] / hh)) + 1 * (0.5 * (-(kap)*h_1 / hxhxh));
                } else {
                  sa[s] = 1 * (0.5 * (-vt[c] / hh)) + 1 * (0.5 * (-(kap)*hh / hxhxh));
                  sb[s] = 1 / tau + 1 * (0.5 * (vt[c] / hh)) + 1 * (0.5 * (kap)*hx2 / hxhxh);
                  scc[s] = 1 * (0.5 * (-(kap)*h_1 / hxhxh));
                }
              }
              double F = (1 / tau) * f_prev[c] + rp[c];
              if (scheme == 0) F = F + 1 * (0.5 * (-vt[c]) * (f_prev[c + sa_] - f_prev[c - sa_]) / hx2);
              else {
                if (vt[c] < 0) F = F + 1 * (0.5 * (-vt[c]) * (f_prev[c + sa_] - f_prev[c]) / hh);
                else F = F + 1 * (0.5 * (-vt[c]) * (f_prev[c] - f_prev[c - sa_]) / hh);
              }
              F = F + 1 * (0.5 * (kap) * ((hh * f_prev[c - sa_] - hx2 * f_prev[c] + h_1 * f_prev[c + sa_]) / hxhxh));
              /* the two other directions in increasing order */
              for (int e = 0; e < 3; e++) {
                if (e == a) continue;
                int de = e + 1;
                const double *fd = de > dir ? f_prev : (de == 1 ? stage[0] : stage[1]);
                double FC, FD;
                side_terms(g, e, c, idx3[e], fd, f_prev, vels[e], kap, scheme, &FC, &FD);
                F = F + 1 * FC + 1 * FD;
              }
              srp[s] = F;
            }
            if (st != XO_OK) break;
            xo_thomas(sd, sv, sa, sb, scc, srp);
            double *fw = stage[dir - 1];
            for (int s = fl; s < sd + 1 - fr; s++) fw[base + (int64_t)(beg + s - fl) * sa_] = sv[s];
            beg = -1;
          }
        }
    }
  }
  free(stage[0]);
  free(stage[1]);
  free(sv); free(sa); free(sb); free(scc); free(srp);
  return st;
}

/* ----------------------------------------------------- immersed boundary */

/* phi_h(r) = (1 + cos(pi r / 2h)) / (4h) for r < 2h: XProblem.cpp:9008-9032. */
double xo_ib_delta(double r, double h) {
  if (r < 2 * h) return 1.0 / (4 * h) * (1 + cos(M_PI * r / (2 * h)));
  return 0;
}

/* Staggered-aware base index, clamped to [3, end_index-1-3]:
 * XProblem.cpp:8931-8967 with range_down = range_up = 3. */
int xo_ib_index(double coord, int type, int axis, double h, int end_index) {
  int index = (int)floor((coord + (type != axis ? 0.5 * h : 0)) / h);
  if (index > (end_index - 1) - 3) index = (end_index - 1) - 3;
  if (index < 3) index = 3;
  return index;
}

static double ib_coord(int index, int type, int axis, double h) {
  return index * h - (type != axis ? 0.5 * h : 0); /* XProblem.cpp:8988-9006 */
}

/* Interpolate (XProblem.cpp:8620-8791): U_n = sum over the 7^3 window of
 * (vel * wx * wy * wz) * h_axis^3, summed in i->j->k order. */
void xo_ib_interp(const xo_grid *g, int axis, int nn, const double *xyz, const double *vel, double *U) {
  double h[3] = {g->hx[0], g->hy[0], g->hz[0]};
  int E[3] = {g->n - 1, g->l - 1, g->m - 1};
  double ha = h[axis - 1];
  double cb = ha * ha * ha;
#pragma omp parallel for schedule(static)
  for (int q = 0; q < nn; q++) {
    const double *X = xyz + 3 * q;
    int b0 = xo_ib_index(X[0], 1, axis, h[0], E[0]);
    int b1 = xo_ib_index(X[1], 2, axis, h[1], E[1]);
    int b2 = xo_ib_index(X[2], 3, axis, h[2], E[2]);
    double acc = 0;
    for (int i = b0 - 3; i <= b0 + 3; i++)
      for (int j = b1 - 3; j <= b1 + 3; j++)
        for (int k = b2 - 3; k <= b2 + 3; k++) {
          double wx = xo_ib_delta(fabs(X[0] - ib_coord(i, 1, axis, h[0])), h[0]);
          double wy = xo_ib_delta(fabs(X[1] - ib_coord(j, 2, axis, h[1])), h[1]);
          double wz = xo_ib_delta(fabs(X[2] - ib_coord(k, 3, axis, h[2])), h[2]);
          acc += (vel[IDX(g, i, j, k)] * wx * wy * wz) * cb;
        }
    U[q] = acc;
  }
}

/* SpreadForce (XProblem.cpp:8550-8617) without the mask side effect (which
 * reads an uninitialised node type, XProblem.cpp:8598): force is zeroed, then
 * force += F_n * wx * wy * wz over each node's window, nodes in order. */
void xo_ib_spread(const xo_grid *g, int axis, int nn, const double *xyz, const double *F, double *force) {
  double h[3] = {g->hx[0], g->hy[0], g->hz[0]};
  int E[3] = {g->n - 1, g->l - 1, g->m - 1};
  int64_t nc = (int64_t)g->n * g->l * g->m;
  memset(force, 0, sizeof(double) * nc);
  for (int q = 0; q < nn; q++) {
    const double *X = xyz + 3 * q;
    int b0 = xo_ib_index(X[0], 1, axis, h[0], E[0]);
    int b1 = xo_ib_index(X[1], 2, axis, h[1], E[1]);
    int b2 = xo_ib_index(X[2], 3, axis, h[2], E[2]);
    for (int i = b0 - 3; i <= b0 + 3; i++)
      for (int j = b1 - 3; j <= b1 + 3; j++)
        for (int k = b2 - 3; k <= b2 + 3; k++) {
          double wx = xo_ib_delta(fabs(X[0] - ib_coord(i, 1, axis, h[0])), h[0]);
          double wy = xo_ib_delta(fabs(X[1] - ib_coord(j, 2, axis, h[1])), h[1]);
          double wz = xo_ib_delta(fabs(X[2] - ib_coord(k, 3, axis, h[2])), h[2]);
          force[IDX(g, i, j, k)] += (F[q] * wx * wy * wz);
        }
  }
}

/* ------------------------------------------------------- projection glue */

/* Pressure right-hand side, TVPProblem::SolveBySplitting XProblem.cpp:5474-5513
 * (staggered fields on the pressure array shape). */
int xo_pressure_rhs(int n, int l, int m, const int32_t *codes, const double *u, const double *v, const double *w,
                    const double *rho, const double *hxU, const double *hyV, const double *hzW, const double *pbc,
                    double tau, double *rhs) {
  int64_t sI = (int64_t)l * m, sJ = m;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < l; j++)
      for (int k = 0; k < m; k++) {
        int64_t c = ((int64_t)i * l + j) * m + k;
        int code = codes[c];
        if (code == 0) rhs[c] = 0;
        else if (code == 1) {
          if (i < 1 || j < 1 || k < 1) return XO_EINVAL;
          double d = rho ? rho[c] : 1.0;
          double div_ = d * d * ((u[c] - u[c - sI]) / hxU[i - 1]) + ((v[c] - v[c - sJ]) / hyV[j - 1]) +
                        ((w[c] - w[c - 1]) / hzW[k - 1]);
          rhs[c] = -(div_ / tau);
        } else if (code == 4) rhs[c] = pbc ? pbc[c] : 0.0;
        else return XO_EINVAL;
      }
  return XO_OK;
}

/* Velocity correction, XProblem.cpp:5598-5778 (axis 1..3). */
int xo_velocity_correct(int n, int l, int m, int axis, const int32_t *codes, const double *us, const double *p,
                        const double *rho, const double *hP, const double *bc, double tau, double *out) {
  int64_t S[3] = {(int64_t)l * m, m, 1};
  int E[3] = {n, l, m};
  int a = axis - 1;
  for (int pass = 1; pass <= 2; pass++)
    for (int i = 0; i < n; i++)
      for (int j = 0; j < l; j++)
        for (int k = 0; k < m; k++) {
          int64_t c = ((int64_t)i * l + j) * m + k;
          int code = codes[c];
          if ((pass == 1) == (code == 4)) continue;
          if (code == 0) out[c] = 0;
          else if (code == 1 || code == 5 || code == 7) {
            int idx[3] = {i, j, k};
            if (idx[a] + 1 >= E[a]) return XO_EINVAL;
            double gradP = (p[c + S[a]] - p[c]) / hP[idx[a]];
            out[c] = us[c] - tau * gradP / (rho ? rho[c] : 1.0);
          } else if (code == 3) {
            double dv = bc ? bc[c] : 0.0;
            if (us[c] != dv) return XO_EINVAL;
            out[c] = dv;
          } else if (code == 4) {
            int64_t nb = -1;
            int cnt = 0;
            for (int dir = 0; dir < 3; dir++)
              for (int sign = 0; sign < 2; sign++) {
                int sh = -1 + sign * 2;
                int aa = i + (dir == 0 ? sh : 0), bb = j + (dir == 1 ? sh : 0), ee = k + (dir == 2 ? sh : 0);
                if (aa >= 0 && aa < n && bb >= 0 && bb < l && ee >= 0 && ee < m) {
                  int64_t cc = ((int64_t)aa * l + bb) * m + ee;
                  if (codes[cc] == 1) { nb = cc; cnt++; }
                }
              }
            if (cnt != 1) return XO_EINVAL;
            double dv = bc ? bc[c] : 0.0;
            if (fabs(us[c] - (2 * dv - us[nb])) > 1e-10) return XO_EINVAL;
            out[c] = 2 * dv - out[nb];
          } else return XO_EINVAL;
        }
  return XO_OK;
}

/* Control balance, XProblem.cpp:5782-5830: max |div u| over Actual cells and
 * the first cell (i->j->k order) attaining it (-1 if all are zero). */
double xo_divergence_max(int n, int l, int m, const int32_t *codes, const double *u, const double *v, const double *w,
                         const double *hxU, const double *hyV, const double *hzW, int64_t *cell) {
  int64_t sI = (int64_t)l * m, sJ = m, best_c = -1;
  double best = 0;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < l; j++)
      for (int k = 0; k < m; k++) {
        int64_t c = ((int64_t)i * l + j) * m + k;
        if (codes[c] != 1 || i < 1 || j < 1 || k < 1) continue;
        double div = ((u[c] - u[c - sI]) / hxU[i - 1]) + ((v[c] - v[c - sJ]) / hyV[j - 1]) +
                     ((w[c] - w[c - 1]) / hzW[k - 1]);
        if (fabs(div) > best) { best = fabs(div); best_c = c; }
      }
  if (cell) *cell = best_c;
  return best;
}

/* Transport velocities, viscosity and boundary right-hand side of one
 * transport equation, as TVPProblem::SolveBySplitting builds them before the
 * LOD step.  axis 1..3 (velocity component eq, on that component's grid):
 * the point-type loop XProblem.cpp:5090-5303 and the viscosity-gradient term
 * at Actual points XProblem.cpp:5357-5417 (with vis; the external force the
 * reference adds at Actual points first, XProblem.cpp:5321-5349, is the
 * caller's).  axis 0 (concentration, on its grid): XProblem.cpp:5889-5944.
 * Staggered fields on the pressure array shape; vis at the pressure nodes
 * (NULL: tvis and the gradient term skipped); bc the boundary values (NULL:
 * 0); tvis and rp may be NULL.  Returns XO_EINVAL where the reference throws
 * (and where an Actual point's stencil would leave the array). */
int xo_transport_velocity(int n, int l, int m, int axis, const int32_t *codes, const double *normals,
                          const double *u, const double *v, const double *w, const double *vis, const double *bc,
                          const double *hx, const double *hy, const double *hz, double *tu, double *tv, double *tw,
                          double *tvis, double *rp) {
  const int64_t S[3] = {(int64_t)l * m, m, 1};
  const int E[3] = {n, l, m};
  const double *f[3] = {u, v, w};
  double *o[3] = {tu, tv, tw};
  const int a = axis - 1;
  if (axis < 0 || axis > 3) return XO_EINVAL;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < l; j++)
      for (int k = 0; k < m; k++) {
        const int64_t c = ((int64_t)i * l + j) * m + k;
        const int idx[3] = {i, j, k};
        const int code = codes[c];
        const double bcv = bc ? bc[c] : 0.0;
        if (axis == 0) {
          /* XProblem.cpp:5896-5940 */
          if (code == 0) {
            tu[c] = 0; tv[c] = 0; tw[c] = 0;
            if (rp) rp[c] = 0;
          } else if (code == 1) {
            if (i < 1 || j < 1 || k < 1) return XO_EINVAL;
            double uu = (u[c - S[0]] + u[c]) / 2;
            tu[c] = uu;
            double vv = (v[c - S[1]] + v[c]) / 2;
            tv[c] = vv;
            double ww = (w[c - S[2]] + w[c]) / 2;
            tw[c] = ww;
            if (rp) rp[c] = 0;
          } else if (code == 4) {
            tu[c] = 0; tv[c] = 0; tw[c] = 0;
            if (rp) rp[c] = i == 0 ? bcv : 0; /* the inlet condition on i = 0 only */
          } else if (code == 6) {
            tu[c] = 0; tv[c] = 0; tw[c] = 0;
            if (rp) rp[c] = 0;
          } else {
            return XO_EINVAL;
          }
          continue;
        }
        if (code == 0) {
          tu[c] = 0; tv[c] = 0; tw[c] = 0;
          if (rp) rp[c] = 0;
          if (tvis) tvis[c] = 0;
        } else if (code == 1) {
          /* the component itself; each other one over the four faces around
           * the node, in the reference's order (c, +a, -e, +a-e) */
          if (idx[a] + 1 >= E[a]) return XO_EINVAL;
          for (int e = 0; e < 3; e++) {
            if (e == a) {
              o[e][c] = f[e][c];
              continue;
            }
            if (idx[e] < 1) return XO_EINVAL;
            const double *fe = f[e];
            o[e][c] = (fe[c] + fe[c + S[a]] + fe[c - S[e]] + fe[c + S[a] - S[e]]) / 4;
          }
          if (tvis && vis) tvis[c] = (vis[c] + vis[c + S[a]]) / 2;
          if (rp) rp[c] = 0;
        } else if (code == 3 || code == 4 || code == 5) {
          if (code == 5) {
            /* Normal points: only for eq 1 with the normal along x (XProblem.cpp:5163-5197) */
            if (axis != 1 || !normals) return XO_EINVAL;
            const double nx = normals[3 * c], ny = normals[3 * c + 1], nz = normals[3 * c + 2];
            if ((nx == 0) || (ny != 0) || (nz != 0)) return XO_EINVAL;
          }
          tu[c] = 0; tv[c] = 0; tw[c] = 0;
          if (tvis) tvis[c] = 0;
          if (rp) rp[c] = bcv;
        } else {
          return XO_EINVAL;
        }
      }
  if (axis == 0 || !vis || !tvis || !rp) return XO_OK;
  /* -nabla sigma at Actual points (XProblem.cpp:5357-5417) */
  const double *tr[3] = {tu, tv, tw};
  for (int i = 0; i < n; i++)
    for (int j = 0; j < l; j++)
      for (int k = 0; k < m; k++) {
        const int64_t c = ((int64_t)i * l + j) * m + k;
        if (codes[c] != 1) continue;
        if (i < 1 || j < 1 || k < 1 || i + 1 >= n || j + 1 >= l || k + 1 >= m) return XO_EINVAL;
        double dVisX = (tvis[c + S[0]] - tvis[c - S[0]]) / (hx[i] + hx[i - 1]);
        double dVisY = (tvis[c + S[1]] - tvis[c - S[1]]) / (hy[j] + hy[j - 1]);
        double dVisZ = (tvis[c + S[2]] - tvis[c - S[2]]) / (hz[k] + hz[k - 1]);
        /* the derivatives of the three transport velocities along the equation's axis */
        const double *ha = a == 0 ? hx : (a == 1 ? hy : hz);
        const int ia = a == 0 ? i : (a == 1 ? j : k);
        double dU = (tu[c + S[a]] - tu[c - S[a]]) / (ha[ia] + ha[ia - 1]);
        double dV = (tv[c + S[a]] - tv[c - S[a]]) / (ha[ia] + ha[ia - 1]);
        double dW = (tw[c + S[a]] - tw[c - S[a]]) / (ha[ia] + ha[ia - 1]);
        const double *velTr = tr[a];
        double dVelX = (velTr[c + S[0]] - velTr[c - S[0]]) / (hx[i] + hx[i - 1]);
        double dVelY = (velTr[c + S[1]] - velTr[c - S[1]]) / (hy[j] + hy[j - 1]);
        double dVelZ = (velTr[c + S[2]] - velTr[c - S[2]]) / (hz[k] + hz[k - 1]);
        rp[c] = rp[c] + (dVisX * (dVelX + dU) + dVisY * (dVelY + dV) + dVisZ * (dVelZ + dW));
      }
  return XO_OK;
}
