// Internal declarations shared by the HIP translation units of libxflow_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "../../include/xflow.h"

// ------------------------------------------------------------------ errors
struct XfError {
  int code;
  std::string msg;
};

#define XF_HIP(call)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      throw XfError{XF_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)};       \
  } while (0)

#define XF_REQUIRE(cond, code, msg)                  \
  do {                                               \
    if (!(cond)) throw XfError{(code), (msg)};       \
  } while (0)

// ------------------------------------------------------------ constants
constexpr int kBlock = 256;          // threads per block for 1-D kernels
constexpr int kMaxPartials = 8192;   // reduction partials per reduction slot
constexpr int kSlots = 8;            // concurrent reductions per kernel

// Point types (XMask.h:6-7, XGrid.h:11, XOperator.h:13-18).
enum : uint8_t {
  PT_FICTIVE = 0, PT_ACTUAL = 1, PT_BORDER = 2, PT_DEFINED = 3, PT_PREDEFINED = 4,
  PT_NORMAL = 5, PT_PRENORMAL = 6, PT_EQUATION = 7, PT_PREEQUATION = 8
};

// Per-cell flag byte of the throughput stencil.
enum : uint8_t {
  FL_IM = 1, FL_IP = 2, FL_JM = 4, FL_JP = 8, FL_KM = 16, FL_KP = 32,  // face open (neighbour not Fictive)
  FL_ACTUAL = 64,                                                     // Actual row
  FL_UNKNOWN = 128,                                                   // space mask 1
};

// Device-resident solver scalars.  Slots are named per solver.
struct DevState {
  double s[32];
  int iter;
  int done;
  int status;
  int pad;
  unsigned counter[kSlots];
};

// Grow-only device workspaces of a context, one per use; every call that
// uses one runs on the context's stream, so a later call's reuse is ordered
// after an earlier call's kernels.
enum WsSlot {
  WS_LOD_ST0, WS_LOD_ST1, WS_LOD_ROWS, WS_LOD_SCRA,
  WS_IB_WIN, WS_IB_BOX, WS_IB_PART,
  WS_THOMAS_A, WS_THOMAS_B,
  WS_BT_VECS, WS_BT_PART,
  WS_COUNT
};

// Errors that only the data can show (a boundary value that does not match
// the predictor): kernels OR these bits into the context's error word; the
// call reads it back, or with deferred errors on, xf_check_errors does.
enum ErrBit {
  EB_VC_DEFINED = 1,     // velocity correction: u* differs from the Defined value
  EB_VC_PREDEF = 2,      // velocity correction: u* at a PreDefined point is not the mirror
};
const char *err_bit_message(int bits);

struct xf_ctx_s {
  int device = 0;
  void *ws[WS_COUNT] = {};
  size_t ws_cap[WS_COUNT] = {};
  int *d_err = nullptr;           // ErrBit word
  bool defer_errors = false;
  hipStream_t stream = nullptr;
  hipStream_t own_stream = nullptr;
  std::string err;
  double *partials = nullptr;     // [kSlots][kMaxPartials]
  unsigned *counters = nullptr;   // [kSlots] last-block counters (self-resetting)
  double *d_scalar = nullptr;     // [kSlots] results of standalone reductions
  int32_t *d_iscalar = nullptr;   // [kSlots]
  DevState *state = nullptr;      // solver state
  DevState *h_state = nullptr;    // pinned mirror
  DevState *h_state2 = nullptr;   // a second pinned mirror (the one-batch-late check)
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t evp[2] = {nullptr, nullptr};  // the state copies of the one-batch-late check
  // per-kernel event timing of the throughput solvers (xf_profile_enable)
  bool prof = false;
  std::vector<hipEvent_t> prof_events;
  std::map<std::string, std::pair<double, long long>> prof_acc;  // name -> (total ms, launches)
};
// Event-bracketed launches: prof_begin() returns an index into a pool of
// (start, stop) event pairs; prof_collect() adds the elapsed times of the
// first `executed` pairs of each name after the stream has synchronised.
struct ProfMark {
  const char *name;
  int ev;
};
int prof_begin(xf_ctx c);
void prof_end(xf_ctx c, int ev);
void prof_collect(xf_ctx c, std::vector<ProfMark> &marks, const std::map<std::string, long long> &executed);

struct xf_grid_s {
  int n = 0, l = 0, m = 0;
  int64_t ncell = 0, nunk = 0;
  std::vector<double> hx, hy, hz;     // host copies of the steps
  double *d_hx = nullptr, *d_hy = nullptr, *d_hz = nullptr;
  // throughput-stencil coefficients: c?[2*i] = 1/(h[i-1]*hh), c?[2*i+1] = 1/(h[i]*hh),
  // hh = (h[i-1]+h[i])/2; zero at the array ends
  double *d_cx = nullptr, *d_cy = nullptr, *d_cz = nullptr;
  float *d_cxf = nullptr, *d_cyf = nullptr, *d_czf = nullptr;
  uint8_t *d_codes = nullptr;
  uint8_t *d_flags = nullptr;
  uint8_t *d_space = nullptr;
  int32_t *d_rn = nullptr;
  double *d_normals = nullptr;        // [ncell*3] or null
  uint8_t *d_has_normal = nullptr;    // [ncell] or null
  double *d_rho = nullptr;            // divergented operator density (owned copy)
  bool use_rho = false;
  int64_t bad_rows_lap = 0;           // rows the Laplace operator rejects
  int64_t bad_rows_div = 0;           // rows the divergented operator rejects
  int64_t special_rows = 0;           // unknown rows that are not Actual (codes 4, 5, 6, 8)
  int ring_unknowns = -1;             // unknowns on the array faces (-1: not counted yet)
  int64_t *d_other_rows = nullptr;    // cells of the unknown rows that are not Actual (fused BiCGStab)
  int64_t n_other_rows = -1;          // their count (-1: not listed yet)
  // xf_transport_velocity per equation (0 concentration, 1..3 velocity
  // components): why the reference would throw on this grid's point types
  // (empty: it would not), and whether every Actual point has both
  // neighbours on each axis (the viscosity-gradient term reads them)
  std::string tv_err[4];
  bool tv_visterm_ok[4] = {false, false, false, false};
  struct CgWork *work = nullptr;      // throughput-solver vectors, allocated on first use
  struct MgHier *mg = nullptr;        // multigrid hierarchy (XF_FCG_MGNN), built on first use
  // LOD transport: sector check of the point types (done on first use; 0 =
  // every line valid, else the failing rule's bit) and, for lines along k,
  // the point types in line-interleaved order (type of cell k of line t at
  // k * (n * l) + t)
  bool lod_checked = false;
  int lod_err = 0;
  uint8_t *d_codes_kt = nullptr;
  int32_t *d_lod_border = nullptr;  // cells neither Actual nor Fictive (the LOD's border passes)
  int64_t n_lod_border = 0;
  int32_t *d_predef = nullptr;  // PreDefined cells (velocity correction)
  int64_t n_predef = -1;
  // projection glue: point-type errors per call form (computed on first use,
  // -1 = not yet), and device copies of the face steps last passed
  int glue_rhs_err = -1;
  int glue_vc_err[3] = {-1, -1, -1};
  std::vector<double> steps_h[4];     // hxU, hyV, hzW, and the velocity correction's hP
  double *steps_d[4] = {nullptr, nullptr, nullptr, nullptr};
  std::vector<uint8_t> codes_h;       // host copy of the point types, made on first need
};
void *ws_get(xf_ctx c, int slot, size_t bytes);
void raise_device_errors(xf_ctx c);  // unless errors are deferred: read, clear and raise the error word
const std::vector<uint8_t> &host_codes(xf_grid g);
const double *device_steps(xf_ctx c, xf_grid g, int slot, const double *h, int cnt);
void free_cg_work(struct CgWork *w);

// T1dPackOperator on the device (xf_pack.hip)
struct xf_pack_s {
  xf_grid g = nullptr;
  int64_t nunk = 0;
  int64_t *cell = nullptr;               // [nunk] array cell of each Rn index
  int32_t *ptr[2] = {nullptr, nullptr};  // [nunk + 1] row starts of A, A^T
  int32_t *col[2] = {nullptr, nullptr};
  double *val[2] = {nullptr, nullptr};
  int64_t nnz[2] = {0, 0};
};
void launch_pack_apply(xf_ctx c, xf_pack_s *p, int transpose, const double *x, const double *b, double sign_b,
                       double *y, bool gated);
void launch_pack_apply_dot(xf_ctx c, xf_pack_s *p, int transpose, const double *x, const double *w, double *out);
void free_mg(struct MgHier *h);
void free_grid(xf_grid g);

// ---------------------------------------------------------------- helpers
inline int grid_blocks(int64_t n, int bs = kBlock) {
  int64_t b = (n + bs - 1) / bs;
  return (int)(b > 0 ? b : 1);
}

// A vector store that streams past the caches (non-temporal) where the
// vector is a whole number of 16-byte chunks, for fields that only a later
// launch reads again.
template <typename T>
__device__ inline void store_stream(T *dst, const void *src, int bytes) {
  if (bytes % 16 == 0) {
    typedef float q4 __attribute__((ext_vector_type(4)));
    for (int h = 0; h < bytes / 16; h++) {
      q4 t;
      __builtin_memcpy(&t, (const char *)src + 16 * h, 16);
      __builtin_nontemporal_store(t, (q4 *)((char *)dst + 16 * h));
    }
  } else {
    __builtin_memcpy(dst, src, bytes);
  }
}

// Sum of v over the block; valid in thread 0.  blockDim.x must be a multiple of 64.
__device__ inline double block_sum(double v, double *lds) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  double r = 0;
  if (threadIdx.x == 0) {
    int nw = blockDim.x >> 6;
    for (int w = 0; w < nw; w++) r += lds[w];
  }
  __syncthreads();
  return r;
}

// Deterministic grid-wide sum: every block deposits its partial; the last block
// to arrive adds the partials in block order and returns true in all its
// threads with the total in *total (thread 0).  The counter resets itself.
__device__ inline bool grid_sum_last(double blockval, double *partials, unsigned *counter, double *total,
                                     double *lds) {
  __shared__ bool am_last;
  int nb = gridDim.x * gridDim.y * gridDim.z;
  int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  if (threadIdx.x == 0) {
    partials[bid] = blockval;
    __threadfence();
    unsigned prev = atomicAdd(counter, 1u);
    am_last = (prev == (unsigned)nb - 1);
  }
  __syncthreads();
  if (!am_last) return false;
  __threadfence();
  double v = 0;
  for (int q = threadIdx.x; q < nb; q += blockDim.x)
    v += __hip_atomic_load(&partials[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  double t = block_sum(v, lds);
  if (threadIdx.x == 0) {
    *total = t;
    *counter = 0;
  }
  return true;
}

// entry-point wrapper
template <class F>
int guarded(xf_ctx ctx, F &&f) {
  try {
    f();
    return XF_OK;
  } catch (const XfError &e) {
    if (ctx) ctx->err = e.msg;
    return e.code;
  } catch (const std::exception &e) {
    if (ctx) ctx->err = e.what();
    return XF_EINVAL;
  }
}

// Internal launchers (defined in the .hip files).
void launch_apply_exact(xf_ctx ctx, xf_grid g, const double *x, const double *b, double sign_b, double *y);
double run_apply_dot_exact(xf_ctx ctx, xf_grid g, const double *x, const double *w, int slot,
                           double *dev_out);
void launch_lincomb(xf_ctx ctx, xf_grid g, double p1, const double *s1, double p2, const double *s2,
                    double *res);
double run_dot(xf_ctx ctx, xf_grid g, const double *a, const double *b, int slot, double *dev_out);
void check_kernel(const char *what);
void launch_apply_gated(xf_ctx ctx, xf_grid g, const double *x, const double *b, double sign_b, double *y);
void launch_apply_dot_gated(xf_ctx ctx, xf_grid g, const double *x, const double *w, double *out, int rslot);
// throughput CG family (xf_cg.hip)
void solve_hs_cg(xf_ctx ctx, xf_grid g, const double *b, const double *x0, double *x, const xf_solve_opts &o,
                 xf_solve_stats &s);
void solve_bicgstab_t(xf_ctx c, xf_grid g, const double *b, const double *x0, double *x, const xf_solve_opts &o,
                      xf_solve_stats &s);
void solve_mp_cg(xf_ctx ctx, xf_grid g, const double *b, const double *x0, double *x, const xf_solve_opts &o,
                 xf_solve_stats &s);
double true_residual(xf_ctx ctx, xf_grid g, const double *b, const double *x, const double *x0, double *work);
// plain FP64 CG on the library's fixed per-device work array (xf_solve_noalloc)
void noalloc_setup(int device);
int64_t noalloc_max_cells();
void solve_hs_cg_noalloc(xf_ctx ctx, xf_grid g, const double *b, const double *x0, double *x,
                         const xf_solve_opts &o, xf_solve_stats &s, double **verify_tmp);
void solve_fcg_nn(xf_ctx ctx, xf_grid g, const double *b, const double *x0, double *x, const xf_solve_opts &o,
                  xf_solve_stats &s);
// the learned preconditioner z = M(r) on a grid (xf_unet.hip); d_sumsq = sum of r^2 over the unknowns
void unet_apply_grid(xf_ctx c, struct xf_unet_s *u, xf_grid g, const float *r32, float *z32, const double *d_sumsq);
void mg_apply(xf_ctx c, xf_grid g, struct xf_unet_s *net, const float *r, float *z);
