/*
 * xflow.h — C ABI of libxflow_hip.so, the MI355X (gfx950) projection-step hot path.
 *
 * Plain C: opaque handles, device pointers, sizes, status codes.  Every entry
 * point returns 0 on success or a negative XF_E* code; xf_last_error() gives
 * the message.  The C++ drop-in classes (xflow_amd/host) turn a non-zero code
 * into `throw "<message>"`, the reference's error convention (XFlow.cpp:78-84).
 *
 * Data layout.  A grid has n x l x m nodes per axis (n = EndIndex_x + 1 ...,
 * the array extents of the reference's T3dNumberMask, XMask.h:90-131).  Every
 * field is one flat device array of n*l*m elements with k fastest,
 * idx = (i*l + j)*m + k: the reference's [i][j][k] order.  Unknowns are the
 * points whose space mask is 1 (XProblem.cpp:4155-4182); vector operations
 * act on unknowns only and leave other entries untouched, as the reference's
 * masked loops do (XSpace.cpp:2293-3487).
 *
 * Threading: one xf_ctx per host thread / HIP stream.  Calls on one context
 * are stream-ordered; calls that return a host scalar synchronise the stream.
 */
#ifndef XFLOW_H
#define XFLOW_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XF_OK 0
#define XF_EINVAL -1      /* bad argument, or the reference would throw     */
#define XF_EMAXIT -2      /* iteration cap reached before the tolerance      */
#define XF_EBREAKDOWN -3  /* Krylov breakdown (zero denominator)            */
#define XF_ENAN -4        /* NaN/Inf residual                                */
#define XF_EHIP -5        /* HIP runtime error                              */
#define XF_ENOMEM -6      /* device allocation failed                        */
#define XF_ESTALL -7      /* preconditioned solve stalled (fallback taken)   */

typedef struct xf_ctx_s *xf_ctx;
typedef struct xf_grid_s *xf_grid;

/* ------------------------------------------------------------- context */
/* No reference counterpart: the reference is single-process CPU code. */
int xf_init(int device, xf_ctx *out);
int xf_shutdown(xf_ctx ctx);
const char *xf_last_error(xf_ctx ctx);
int xf_set_stream(xf_ctx ctx, void *hip_stream); /* NULL: the context's own stream */
int xf_synchronize(xf_ctx ctx);
/* Checks that only the data can fail (xf_velocity_correct's boundary-value
 * checks, where the reference throws) are read back by the call itself by
 * default, which waits for the stream.  With deferred errors on, calls only
 * record them on the device; xf_check_errors waits for the stream, returns
 * XF_EINVAL with the first recorded error's message (xf_last_error) and
 * clears the record. */
int xf_set_deferred_errors(xf_ctx ctx, int on);
int xf_check_errors(xf_ctx ctx);
int xf_malloc(xf_ctx ctx, size_t bytes, void **dptr);
int xf_free(xf_ctx ctx, void *dptr);
/* kind: 0 host->device, 1 device->host, 2 device->device; synchronous. */
int xf_memcpy(xf_ctx ctx, void *dst, const void *src, size_t bytes, int kind);

/* ---------------------------------------------------------------- grid */
/* Replaces T3dNormalGrid(sep1, sep2, sep3, mask, normals) (XGrid.h:282,
 * XGrid.cpp:414-441) together with TRnRelease3dSpaceModel(1, &spaceMask)
 * (XSpaceModel.h:502, XSpaceModel.cpp:276-411) built from the space-mask
 * rule of XProblem.cpp:4155-4182.  hx[n-1], hy[l-1], hz[m-1] are the
 * separator steps (TSeparator::Separation, XSeparator.cpp:25-99); codes are
 * the point types 0..8; normals (n*l*m*3, outward, axis-aligned where the
 * operator needs them) and has_normal may be NULL.  Host pointers. */
int xf_grid_create(xf_ctx ctx, int n, int l, int m, const double *hx, const double *hy, const double *hz,
                   const int32_t *codes, const double *normals, const int32_t *has_normal, xf_grid *out);
int xf_grid_destroy(xf_ctx ctx, xf_grid g);
/* ncell = n*l*m; n_unknowns = TRnSpaceModel::EndIndex (XSpaceModel.cpp:330-358). */
int xf_grid_info(xf_grid g, int64_t *ncell, int64_t *n_unknowns);
/* Parity dumps computed on the device (host buffers, each may be NULL):
 * space mask [ncell], Rn index [ncell] (1-based, 0 = not an unknown,
 * XSpaceModel.cpp:381-406), activity table [n_unknowns*8]
 * (T3dLaplaceOperator::ConstructActivityTables, XOperator.cpp:837-936). */
int xf_grid_export(xf_ctx ctx, xf_grid g, int32_t *space_mask, int32_t *rn, int32_t *act);
/* Switches the grid's operator to T3dVariableDivergentedOperator with the
 * given density (device array [ncell]); NULL switches back to
 * T3dLaplaceOperator.  Replaces withDensity (XOperator.cpp:2235-2268); the
 * density is copied into a buffer the grid reuses (no per-call leak). */
int xf_grid_set_density(xf_ctx ctx, xf_grid g, const double *rho);

/* ----------------------------------------------- operator and vectors */
/* y = A x + sign_b * b at unknowns (sign_b in {0, +1, -1}; b may be NULL
 * when sign_b = 0).  Rows follow T3dLaplaceOperator::operator()
 * (XOperator.cpp:938-1129) or the divergented operator
 * (XOperator.cpp:1901-2233), in the reference's arithmetic order; rows the
 * reference rejects give XF_EINVAL.  Replaces `y |= A(x) - b`
 * (XSpace.cpp:3013-3217). */
int xf_apply(xf_ctx ctx, xf_grid g, const double *x, const double *b, double sign_b, double *y);
/* (A x, x) without storing A x (XSpace.cpp:3344-3419); *out is a host scalar. */
int xf_apply_dot(xf_ctx ctx, xf_grid g, const double *x, double *out);
/* res = p1*s1 + p2*s2 at unknowns (XSpace.cpp:2489-3010). */
int xf_lincomb(xf_ctx ctx, xf_grid g, double p1, const double *s1, double p2, const double *s2, double *res);
/* res |= s at unknowns (XSpace.cpp:2293-2363). */
int xf_copy(xf_ctx ctx, xf_grid g, const double *s, double *res);
/* (a, b) over unknowns (XSpace.cpp:3261-3342); host scalar. */
int xf_dot(xf_ctx ctx, xf_grid g, const double *a, const double *b, double *out);
/* max |x| over unknowns and the Rn index of its first occurrence
 * (TRnSpace::GetAbsMaxElement, XSpace.cpp:961-973, 3428-3487). */
int xf_absmax(xf_ctx ctx, xf_grid g, const double *x, double *val, int32_t *rn_index);

/* --------------------------------------------------------------- solvers */
#define XF_CG3 0       /* TSLAE::SolveByConjugateGradients, XSLAE.cpp:56-157      */
#define XF_BICG 1      /* TSLAE::SolveByBiConjugateGradients, XSLAE.cpp:159-348   */
#define XF_MR 2        /* TSLAE::SolveByMinimalResiduals, XSLAE.cpp:350-414      */
#define XF_BICGSTAB 3  /* TSLAE::SolveByBiConjugateGradientsStab, 667-767        */
#define XF_CG 4        /* single-reduction CG, two fused kernels per iteration
                          (throughput path, no reference counterpart)           */
#define XF_CG_MP 5     /* FP32 inner CG + FP64 iterative refinement              */
#define XF_FCG_NN 6    /* flexible CG with the learned preconditioner             */
#define XF_FCG_MGNN 7  /* flexible CG, V-cycle preconditioner with the learned
                          network as its coarsest-grid solve (precond may be
                          NULL: the full-depth classical cycle instead)          */
#define XF_BICGSTAB_T 8 /* the reference's BiCGStab recurrence and stopping rule
                          (max |r| <= tol, absolute; default 1e-5) on its
                          operator rows in four fused passes per iteration       */

#define XF_TOL_ABS 0   /* stop on ||r|| <= tol (the reference's rule)            */
#define XF_TOL_REL 1   /* stop on ||r|| <= tol * ||b - A x0||                      */

typedef struct {
  int method;        /* XF_CG3 ...                                             */
  int tol_mode;      /* XF_TOL_ABS or XF_TOL_REL                                */
  double tol;        /* the reference's hard-coded values: CG 1e-5, BiCG 1e-8,
                        MR 1e-7 (2-norm), BiCGStab 1e-5 (max-norm),
                        XSLAE.cpp:64, 166, 358, 669                              */
  int maxit;         /* iteration cap (the reference has none)                   */
  int check_every;   /* throughput methods: host convergence poll period        */
  int project_mean;  /* remove the mean over unknowns (pure-Neumann systems)     */
  double inner_tol;  /* XF_CG_MP / XF_FCG_NN: relative tolerance of inner sweeps */
  int stall_window;  /* XF_FCG_NN: fall back to plain CG after this many
                        iterations without progress (20)                         */
  void *precond;     /* XF_FCG_NN, XF_FCG_MGNN: an xf_unet handle               */
  int verify;        /* 1 (default): after the solve, recompute
                        ||b - A x|| / ||b - A x0|| with the reference rows into
                        stats.resid_true (two operator passes and a temporary
                        field); 0: skip that check, resid_true = -1               */
} xf_solve_opts;

typedef struct {
  int iters;          /* outer iterations (CG3/BiCG/MR/BiCGStab: the reference's
                         iteration counter)                                     */
  int inner_iters;    /* XF_CG_MP: total FP32 iterations                         */
  int status;         /* XF_OK or an XF_E* code                                  */
  int fallback;       /* XF_FCG_NN: 1 if plain CG finished the solve             */
  double resid;       /* final residual in the method's own stopping norm         */
  double resid_true;  /* ||b - A x||_2 / ||b - A x0||_2 recomputed in FP64       */
  double ms;          /* device time of the solve (HIP events)                   */
} xf_solve_stats;

void xf_solve_opts_default(int method, xf_solve_opts *o);
/* Solves A x = b from x0 (entries of x0 outside the unknowns are kept in x,
 * as the reference's vector copies keep them).  Device pointers. */
int xf_solve(xf_ctx ctx, xf_grid g, const double *b, const double *x0, double *x, const xf_solve_opts *opts,
             xf_solve_stats *stats);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): xf_solve with opts->method XF_CG (the plain FP64 CG,
 * same kernels and same results), its work vectors in one fixed array per
 * device that the library holds from load to process end (72 MiB: grids of
 * at most XF_NOALLOC_MAX_CELLS = n*l*m cells; larger grids give XF_EINVAL).
 * It calls no hipMalloc/hipFree and records no memset or memcpy (state
 * cleared and x0 copied by kernels).  Every call on one GPU shares that
 * array: calls that overlap on one GPU (two streams or threads) race and
 * must be serialised by the caller.  opts->verify uses the same array after
 * the solve.  Per-kernel profiling (xf_profile_enable) is not sampled. */
#define XF_NOALLOC_MAX_CELLS 2358272
int xf_solve_noalloc(xf_ctx ctx, xf_grid g, const double *b, const double *x0, double *x,
                     const xf_solve_opts *opts, xf_solve_stats *stats);

/* ------------------------------------------------------------ transport */
/* One LOD transport step, THydrodynamicsProblem::SolveTransferEquation__
 * (XProblem.cpp:2784-3475; direction order 1, all terms on): f is the field
 * at the new level (updated in place, entries outside the sectors kept),
 * f_prev the previous level, u/v/w the transport velocities, rp the source /
 * boundary values, kappa the diffusion coefficient, scheme 0 central, 1
 * upwind.  Point types come from the grid; sector borders the reference
 * rejects give XF_EINVAL.  Device pointers, all [ncell]. */
int xf_lod_transport(xf_ctx ctx, xf_grid g, double *f, const double *f_prev, const double *u, const double *v,
                     const double *w, const double *rp, const double *kappa, double tau, int scheme);
/* No reference counterpart (the reference advances the velocity components
 * one after another): xf_lod_transport for count (1..4) independent problems,
 * problem p on grids[p] with the fields f[p], f_prev[p], ..., in one set of
 * launches, with the same results as count separate calls.  No problem's
 * output may be another problem's field. */
int xf_lod_transport_batch(xf_ctx ctx, int count, const xf_grid *grids, double *const *f, const double *const *f_prev,
                           const double *const *u, const double *const *v, const double *const *w,
                           const double *const *rp, const double *const *kappa, double tau, int scheme);
/* nsys independent tridiagonal systems a u[m-1] + b u[m] + c u[m+1] = f,
 * m = 1..M-1, u[0] and u[M] given (TSLAE::SolveByScalarRunning,
 * XSLAE.cpp:416-445); system s at [s*(M+1), (s+1)*(M+1)).  Device pointers. */
int xf_thomas_batched(xf_ctx ctx, int nsys, int M, double *u, const double *a, const double *b, const double *c,
                      const double *f);

/* ---------------------------------------------------- immersed boundary */
/* axis 1, 2, 3 = COORD_X, COORD_Y, COORD_Z (XProblem.h:1105-1107); the grid
 * is the velocity component's grid (uniform steps, Separation[0] per axis);
 * xyz[nn*3] node positions, device pointers. */
/* U[q] = h_axis^3 sum_{7^3 window} vel * phi phi phi (Interpolate, XProblem.cpp:8620-8791). */
int xf_ib_interp(xf_ctx ctx, xf_grid g, int axis, int nn, const double *xyz, const double *vel, double *U);
/* force = 0, then force += F[q] phi phi phi over each node's window, nodes in
 * order (SpreadForce, XProblem.cpp:8550-8617, without its mask side effect). */
int xf_ib_spread(xf_ctx ctx, xf_grid g, int axis, int nn, const double *xyz, const double *F, double *force);
/* xyz_prev[.,axis] = xyz[.,axis]; xyz[.,axis] += U dt (UpdateBoundaryPosition,
 * XProblem.cpp:8858-8894); xyz_prev may be NULL. */
int xf_ib_update_position(xf_ctx ctx, int axis, int nn, double *xyz, double *xyz_prev, const double *U, double dt);

/* ------------------------------------------------------- projection glue */
/* Staggered fields live on the pressure array shape: u[i,j,k] is the x-face
 * between pressure cells (i,j,k) and (i+1,j,k) (v, w likewise).  hxU[n-1],
 * hyV[l-1], hzW[m-1] are the velocity grids' separator steps (host arrays).
 * rho, pbc and bc may be NULL (density 1, boundary values 0).  Device fields. */
/* rhs = -(rho^2 du/dx + dv/dy + dw/dz)/tau at Actual cells, 0 at Fictive,
 * pbc at PreDefined (TVPProblem::SolveBySplitting, XProblem.cpp:5474-5513). */
int xf_pressure_rhs(xf_ctx ctx, xf_grid gp, const double *u, const double *v, const double *w, const double *rho,
                    const double *hxU, const double *hyV, const double *hzW, const double *pbc, double tau,
                    double *rhs);
/* u = u* - tau (p[+1]-p)/hP / rho along axis (1..3) on the velocity grid gv,
 * Defined points checked against bc, PreDefined points mirrored about their
 * Actual neighbour (XProblem.cpp:5598-5778); hP = pressure steps (host). */
int xf_velocity_correct(xf_ctx ctx, xf_grid gv, int axis, const double *u_star, const double *p, const double *rho,
                        const double *hP, const double *bc, double tau, double *u_out);
/* max |div u| over Actual pressure cells and its flat cell index (-1 if 0)
 * (control balance, XProblem.cpp:5782-5830). */
int xf_divergence_max(xf_ctx ctx, xf_grid gp, const double *u, const double *v, const double *w, const double *hxU,
                      const double *hyV, const double *hzW, double *max_abs, int64_t *cell);

/* ------------------------------------------------------- packed operator */
/* T1dPackOperator (XOperator.cpp:1135-1571; XOperator.h:323-383): the grid's
 * operator as packed rows over the Rn unknowns, for A and for A^T (operator~),
 * with the rows' entries in activity-table order and zero coefficients left
 * out, as the reference builds them; the product with Rn-ordered vectors sums
 * each row in stored order (T1dPackOperator::operator(), XOperator.cpp:
 * 1551-1568).  Device vectors of length n_unknowns; export copies to host
 * arrays (counts[n_unknowns], vals/cols[nnz], 0-based columns). */
typedef struct xf_pack_s *xf_pack;
int xf_pack_create(xf_ctx ctx, xf_grid g, xf_pack *out);
int xf_pack_destroy(xf_ctx ctx, xf_pack p);
int xf_pack_info(xf_pack p, int64_t *n_unknowns, int64_t *nnz, int64_t *nnz_transpose);
int xf_pack_export(xf_ctx ctx, xf_pack p, int transpose, int32_t *counts, double *vals, int32_t *cols);
int xf_pack_apply(xf_ctx ctx, xf_pack p, int transpose, const double *x, double *y);
/* field [ncell] <-> Rn vector [n_unknowns] (TRnSpace::operator[] order, XSpace.cpp:911-920) */
int xf_pack_gather(xf_ctx ctx, xf_pack p, const double *field, double *vec);
int xf_pack_scatter(xf_ctx ctx, xf_pack p, const double *vec, double *field);
/* TSLAE on the packed operator (XSLAE.cpp:56-414, 667-767 with A a
 * T1dPackOperator): methods XF_CG3, XF_BICG, XF_MR, XF_BICGSTAB on Rn vectors
 * of length n_unknowns.  transpose = 1 solves with A^T (the state operator~
 * leaves the packed operator in, XOperator.h:333); BiCG's shadow sequence uses
 * the other matrix (its A.Copy() then ~, XSLAE.cpp:205-206). */
int xf_pack_solve(xf_ctx ctx, xf_pack p, int transpose, const double *b, const double *x0, double *x,
                  const xf_solve_opts *opts, xf_solve_stats *stats);
/* Rn vectors (TRnRelease1dSpace, every entry an unknown; XSpace.cpp:1823-2125):
 * res = p1*s1 + p2*s2, (a, b), and max |x| with its 1-based index */
int xf_rn_lincomb(xf_ctx ctx, int64_t n, double p1, const double *s1, double p2, const double *s2, double *res);
int xf_rn_dot(xf_ctx ctx, int64_t n, const double *a, const double *b, double *out);
int xf_rn_absmax(xf_ctx ctx, int64_t n, const double *x, double *val, int32_t *rn_index);

/* ------------------------------------------------- learned preconditioner */
/* No reference counterpart: the learned preconditioner of XF_FCG_NN.  A
 * 3-level 3-D U-Net (channels 16/32/64, 3x3x3 convolutions, BF16 weights and
 * activations, FP32 accumulation).  Parameters are one FP32 vector: for each
 * layer c0..c9, W[Cout][27][Cin] (tap = 9 dz + 3 dy + dx, offsets -1..1)
 * then b[Cout].  Crop tensors are device [B][D][H][W] FP32 with D, H
 * multiples of 16 and W a multiple of 64. */
typedef struct xf_unet_s *xf_unet;
int xf_unet_create(xf_ctx ctx, int seed, xf_unet *out);
int xf_unet_destroy(xf_ctx ctx, xf_unet u);
int xf_unet_param_count(xf_unet u, int64_t *count);
int xf_unet_get_params(xf_ctx ctx, xf_unet u, float *host);
int xf_unet_set_params(xf_ctx ctx, xf_unet u, const float *host);
int xf_unet_get_grads(xf_ctx ctx, xf_unet u, float *host);
/* y = net(r, mask) on crops (no scaling) */
int xf_unet_forward(xf_ctx ctx, xf_unet u, int B, int D, int H, int W, const float *r, const float *mask, float *y);
/* the training loss of net(r, mask) against target (xf_unet_set_loss) and its
 * parameter gradient (read with xf_unet_get_grads) */
int xf_unet_loss_grad(xf_ctx ctx, xf_unet u, int B, int D, int H, int W, const float *r, const float *mask,
                      const float *target, double *loss);
/* training loss: XF_LOSS_MSE (default) = mean over the mask of (y - t)^2;
 * XF_LOSS_COSINE = mean over samples of 1 - (y,t)^2/((y,y)(t,t)) over the
 * mask, invariant to the scale of y per sample (what a flexible-CG step sees) */
#define XF_LOSS_MSE 0
#define XF_LOSS_COSINE 1
/* XF_LOSS_ENERGY = mean over samples of 1 - (y,Lt)^2/((y,Ly)(t,Lt)) with L
 * the masked 7-point Laplacian of the crop in cell units (faces closed at the
 * crop edge): the energy-norm error an optimally scaled correction along y
 * leaves of the target t (for whole-domain samples t is the solver error) */
#define XF_LOSS_ENERGY 2
int xf_unet_set_loss(xf_ctx ctx, xf_unet u, int mode);
/* loss_grad followed by one Adam step (beta 0.9 / 0.999, eps 1e-8) */
int xf_unet_train_step(xf_ctx ctx, xf_unet u, int B, int D, int H, int W, const float *r, const float *mask,
                       const float *target, float lr, double *loss);
/* the preconditioner on a pressure grid: z = s net(r/s) on the unknowns, 0
 * elsewhere, s = rms of r over the unknowns (FP32 fields [ncell]) */
int xf_unet_apply(xf_ctx ctx, xf_unet u, xf_grid g, const float *r, float *z);

/* No reference counterpart: the V-cycle of XF_FCG_MGNN on its own.  Grids are
 * halved per axis while the interior extents stay even and the x extent stays
 * at least 64; coarse cells are unknowns if any child is (uniform steps
 * required).  xf_mg_levels builds the hierarchy and returns its depth;
 * xf_mg_apply computes z = M(r) for FP32 fields [ncell] of g (net may be
 * NULL). */
int xf_mg_levels(xf_ctx ctx, xf_grid g, int *levels);
/* the grid of one level (0 = g), owned by g: do not destroy */
int xf_mg_level(xf_ctx ctx, xf_grid g, int level, xf_grid *out);
int xf_mg_apply(xf_ctx ctx, xf_grid g, xf_unet net, const float *r, float *z);

/* The transport terms of one transport equation as TVPProblem::SolveBySplitting
 * builds them before its LOD step (there inline, not a routine of its own):
 * axis 1..3, velocity component eq on that component's grid g (point-type
 * loop XProblem.cpp:5090-5303): at Actual points the component itself and the
 * other two averaged over the four faces around the node in the reference's
 * operand order, tvis = the mean of vis over the two pressure nodes, rp = 0;
 * Fictive points 0; Defined, PreDefined and Normal (eq 1, normal along x)
 * points 0 with rp = bc; with vis, tvis and rp given, rp at Actual points
 * then gains the viscosity-gradient term of XProblem.cpp:5357-5417 (the
 * reference adds its external force first, XProblem.cpp:5321-5349: callers
 * add theirs).  axis 0, the concentration on its grid (XProblem.cpp:5889-5944):
 * cell-centre velocities (each component over its two faces), rp = bc on the
 * PreDefined points of the plane i = 0 (the inlet) and 0 elsewhere; tvis must
 * be NULL.  Point types the reference rejects give XF_EINVAL with its
 * message.  Staggered fields u, v, w and vis (pressure nodes) on g's array
 * shape; vis, bc, tvis, rp may be NULL.  Device pointers. */
int xf_transport_velocity(xf_ctx ctx, xf_grid g, int axis, const double *u, const double *v, const double *w,
                          const double *vis, const double *bc, double *tu, double *tv, double *tw, double *tvis,
                          double *rp);

/* ------------------------------------------------------------ output files */
/* The node-centred ASCII VTK structured grid of TVPProblem::SolveBySplitting
 * (XProblem.cpp:6132-6379), same text: nodes 1..n-2 of each axis (i fastest),
 * velocity = mean of the four faces around a node (0 where mask_grid_u's
 * point type is Fictive), pressure / density / concentration / viscosity =
 * sum of the eight cells around the node over the number of non-Fictive ones
 * among them (masks of mask_grid_p, _d, _c, _c; 0 at Fictive nodes).  Fields
 * are device arrays of the grids' shape; p, density, concentration and
 * viscosity may be NULL (written as 0).  x, y, z: host node coordinates. */
typedef struct {
  xf_grid mask_grid_u, mask_grid_p, mask_grid_d, mask_grid_c;
  const double *x, *y, *z;
  const double *u, *v, *w, *p, *density, *concentration, *viscosity;
} xf_vtk_flow;
int xf_vtk_write_flow(xf_ctx ctx, const char *path, const xf_vtk_flow *fields);
/* The immersed-boundary VTK of TImmersedBoundaryProblem::OutputBoundary
 * (XProblem.cpp:9543-9635), same text: `circles` rings of `in_circle` nodes,
 * positions xyz[3*nodes], start positions xyz0 (displacement = xyz0 - xyz),
 * velocities vel[3*nodes], pressure p[nodes] and wall shear stress
 * wss[nodes], in host or device memory.  `time` is not written (the
 * reference puts it in the file name). */
int xf_vtk_write_boundary(xf_ctx ctx, const char *path, int circles, int in_circle, const double *xyz,
                          const double *xyz0, const double *vel, const double *p, const double *wss, double time);
/* cnpy::npy_save (cnpy.h:66-125, 207-236), same bytes: a C-order array of
 * kind 'f' (word size 4 or 8), 'i' or 'u' (1, 2, 4, 8 bytes) and the given
 * shape, from host or device memory. */
int xf_npy_save(xf_ctx ctx, const char *path, const void *data, char kind, int word_size, int ndims,
                const uint32_t *shape);

/* ------------------------------------------------------------ profiling */
/* No reference counterpart (the reference only prints clock() deltas,
 * XProblem.cpp:5442-5445).  When enabled, the throughput solvers bracket
 * each kernel launch with HIP events on the context's stream; read returns
 * the average device time per launch of a named kernel ("cg_dots_fp64",
 * "cg_step_fp64", "cg_fused_fp64", their "_fp32" forms, "mp_residual_fp64",
 * "mp_correct") since the last reset. */
int xf_profile_enable(xf_ctx ctx, int on);
int xf_profile_reset(xf_ctx ctx);
int xf_profile_read(xf_ctx ctx, const char *kernel, double *avg_ms, int64_t *launches);
/* No reference counterpart: the streaming-bandwidth ceilings of the GPU for
 * the solver kernels' access patterns, measured with two buffers of `bytes`
 * each over `reps` launches after one warm-up launch: kind 0 a read-only
 * reduction, 1 a copy (one read, one write), 2 two reads and two writes in
 * place, each as a grid-stride loop and as a flat launch (plain and
 * non-temporal accesses).  *gbs = bytes moved per second / 1e9 of the
 * fastest form.  Allocates its buffers. */
int xf_stream_ceiling(xf_ctx ctx, int kind, size_t bytes, int reps, double *gbs);

#ifdef __cplusplus
}
#endif
#endif /* XFLOW_H */
