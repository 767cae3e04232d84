/* CPU oracle for the xFlow projection-step hot path — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference algorithms, used by tests/ as the
 * checker, by bench.py as the timed CPU baseline and by __graft_entry__.smoke()
 * as the checker.  The product (libxflow_hip.so) never links or calls it.
 *
 * Conventions (same as the reference's logical layout):
 *   - a grid has n x l x m nodes (n = EndIndex_x + 1, ...), stored flat with
 *     k fastest: idx = (i*l + j)*m + k;
 *   - hx[n-1], hy[l-1], hz[m-1] are the separator step arrays (Separation);
 *   - codes[] are point types 0..8 (XMask.h:6-7, XGrid.h:11, XOperator.h:13-18);
 *   - space_mask[] is 0/1 (unknown or not), Rn indices are 1-based.
 *
 * Parity status of each function is stated in DESIGN.md and in the header of
 * xf_oracle.c; every function is checked against the compiled reference's
 * golden vectors in tests/golden and/or closed forms.
 */
#ifndef XF_ORACLE_H
#define XF_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int n, l, m;
  const double *hx, *hy, *hz;
  const int32_t *codes;       /* point types of the operator grid          */
  const int32_t *space_mask;  /* 0/1 unknowns                              */
  const double *normals;      /* [ncell*3] or NULL                         */
  const int32_t *has_normal;  /* [ncell] or NULL                           */
  const double *rho;          /* NULL: T3dLaplaceOperator; else density of
                                 T3dVariableDivergentedOperator             */
} xo_grid;

/* status codes */
#define XO_OK 0
#define XO_EINVAL -1
#define XO_EMAXIT -2
#define XO_EBREAKDOWN -3
#define XO_ENAN -4

void xo_set_threads(int nthreads);
int  xo_space_mask(int64_t ncell, const int32_t *codes, int32_t *space_mask);
int64_t xo_rn_number(int n, int l, int m, const int32_t *space_mask, int32_t *rn);
int  xo_activity(int n, int l, int m, const int32_t *space_mask, const int32_t *rn, int32_t *act);

int  xo_apply(const xo_grid *g, const double *x, const double *b, double sign_b, double *y);
double xo_apply_dot(const xo_grid *g, const double *x, int *status);
void xo_lincomb(const xo_grid *g, double p1, const double *s1, double p2, const double *s2, double *res);
void xo_copy(const xo_grid *g, const double *s, double *res);
double xo_dot(const xo_grid *g, const double *a, const double *b);
double xo_absmax(const xo_grid *g, const double *x, const int32_t *rn, int32_t *rn_index);

int xo_cg3(const xo_grid *g, const double *b, const double *x0, double *x, double eps, int maxit,
           int *iters, double *resid);
int xo_bicg(const xo_grid *g, const double *b, const double *x0, double *x, double eps, int maxit,
            int *iters, double *resid);
int xo_mr(const xo_grid *g, const double *b, const double *x0, double *x, double eps, int maxit,
          int *iters, double *resid);
int xo_bicgstab(const xo_grid *g, const double *b, const double *x0, double *x, double eps, int maxit,
                int *iters, double *resid);
int xo_cg_hs(const xo_grid *g, const double *b, const double *x0, double *x, double rtol, int maxit,
             int project_mean, int *iters, double *resid);
int xo_cg_cg(const xo_grid *g, const double *b, const double *x0, double *x, double rtol, int maxit,
             int project_mean, int *iters, double *resid);

int64_t xo_pack(const xo_grid *g, int transpose, int32_t *counts, double *vals, int32_t *cols);
void xo_pack_apply(int64_t nunk, const int32_t *counts, const double *vals, const int32_t *cols, const double *x,
                   double *y);
/* method: 0 CG3, 1 BiCG, 2 minimal residuals, 3 BiCGStab */
int xo_pack_solve(int method, int64_t nunk, const int32_t *counts, const double *vals, const int32_t *cols,
                  const int32_t *counts_t, const double *vals_t, const int32_t *cols_t, const double *b,
                  const double *x0, double *x, double eps, int maxit, int *iters, double *resid);

void xo_thomas(int M, double *u, const double *a, const double *b, const double *c, const double *f);
int  xo_lod(const xo_grid *g, double *f, const double *f_prev, const double *u, const double *v,
            const double *w, const double *rp, const double *kappa, double tau, int scheme);

double xo_ib_delta(double r, double h);
int  xo_ib_index(double coord, int type, int axis, double h, int end_index);
void xo_ib_interp(const xo_grid *g, int axis, int nn, const double *xyz, const double *vel, double *U);
void xo_ib_spread(const xo_grid *g, int axis, int nn, const double *xyz, const double *F, double *force);

int xo_pressure_rhs(int n, int l, int m, const int32_t *codes, const double *u, const double *v, const double *w,
                    const double *rho, const double *hxU, const double *hyV, const double *hzW, const double *pbc,
                    double tau, double *rhs);
int xo_velocity_correct(int n, int l, int m, int axis, const int32_t *codes, const double *us, const double *p,
                        const double *rho, const double *hP, const double *bc, double tau, double *out);
double xo_divergence_max(int n, int l, int m, const int32_t *codes, const double *u, const double *v, const double *w,
                         const double *hxU, const double *hyV, const double *hzW, int64_t *cell);
/* Transport velocities, viscosity and boundary right-hand side of one
 * transport equation: XProblem.cpp:5090-5303 and 5357-5417 (axis 1..3),
 * 5889-5944 (axis 0, concentration). */
int xo_transport_velocity(int n, int l, int m, int axis, const int32_t *codes, const double *normals,
                          const double *u, const double *v, const double *w, const double *vis, const double *bc,
                          const double *hx, const double *hy, const double *hz, double *tu, double *tv, double *tw,
                          double *tvis, double *rp);

#ifdef __cplusplus
}
#endif
#endif
