// Context, device memory, grid set-up (point types, space mask, Rn numbering,
// activity table) for libxflow_hip.so.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

#include "xf_internal.h"

void check_kernel(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw XfError{XF_EHIP, std::string(what) + ": " + hipGetErrorString(e)};
}

// ------------------------------------------------------------------ context
extern "C" int xf_init(int device, xf_ctx *out) {
  if (!out) return XF_EINVAL;
  xf_ctx c = new xf_ctx_s();
  int st = guarded(c, [&] {
    int count = 0;
    XF_HIP(hipGetDeviceCount(&count));
    XF_REQUIRE(device >= 0 && device < count, XF_EINVAL, "xf_init: no such device");
    XF_HIP(hipSetDevice(device));
    c->device = device;
    noalloc_setup(device);
    // blocking stream: ordered with the legacy default stream that torch uses
    XF_HIP(hipStreamCreate(&c->own_stream));
    c->stream = c->own_stream;
    XF_HIP(hipMalloc(&c->partials, sizeof(double) * kSlots * kMaxPartials));
    XF_HIP(hipMalloc(&c->counters, sizeof(unsigned) * kSlots));
    XF_HIP(hipMemset(c->counters, 0, sizeof(unsigned) * kSlots));
    XF_HIP(hipMalloc(&c->d_scalar, sizeof(double) * kSlots));
    XF_HIP(hipMalloc(&c->d_iscalar, sizeof(int32_t) * kSlots));
    XF_HIP(hipMalloc(&c->state, sizeof(DevState)));
    XF_HIP(hipMemset(c->state, 0, sizeof(DevState)));
    XF_HIP(hipHostMalloc((void **)&c->h_state, sizeof(DevState), hipHostMallocDefault));
    XF_HIP(hipHostMalloc((void **)&c->h_state2, sizeof(DevState), hipHostMallocDefault));
    XF_HIP(hipEventCreateWithFlags(&c->evp[0], hipEventDisableTiming));
    XF_HIP(hipEventCreateWithFlags(&c->evp[1], hipEventDisableTiming));
    XF_HIP(hipEventCreate(&c->ev0));
    XF_HIP(hipEventCreate(&c->ev1));
    XF_HIP(hipMalloc(&c->d_err, sizeof(int)));
    XF_HIP(hipMemset(c->d_err, 0, sizeof(int)));
  });
  if (st != XF_OK) {
    fprintf(stderr, "xf_init: %s\n", c->err.c_str());
    delete c;
    return st;
  }
  *out = c;
  return XF_OK;
}

extern "C" int xf_shutdown(xf_ctx c) {
  if (!c) return XF_EINVAL;
  hipSetDevice(c->device);
  hipStreamSynchronize(c->stream);
  hipFree(c->partials);
  hipFree(c->counters);
  hipFree(c->d_scalar);
  hipFree(c->d_iscalar);
  hipFree(c->state);
  hipHostFree(c->h_state);
  hipHostFree(c->h_state2);
  if (c->evp[0]) hipEventDestroy(c->evp[0]);
  if (c->evp[1]) hipEventDestroy(c->evp[1]);
  hipEventDestroy(c->ev0);
  hipEventDestroy(c->ev1);
  hipFree(c->d_err);
  for (int k = 0; k < WS_COUNT; k++) hipFree(c->ws[k]);
  for (auto e : c->prof_events) hipEventDestroy(e);
  if (c->own_stream) hipStreamDestroy(c->own_stream);
  delete c;
  return XF_OK;
}

extern "C" const char *xf_last_error(xf_ctx c) { return c ? c->err.c_str() : "null context"; }

extern "C" int xf_set_stream(xf_ctx c, void *s) {
  if (!c) return XF_EINVAL;
  // the workspaces are ordered by the stream: let the old stream's work finish
  if (hipStreamSynchronize(c->stream) != hipSuccess) return XF_EHIP;
  c->stream = s ? (hipStream_t)s : c->own_stream;
  return XF_OK;
}

void *ws_get(xf_ctx c, int slot, size_t bytes) {
  if (bytes == 0) bytes = 1;
  if (c->ws_cap[slot] < bytes) {
    if (c->ws[slot]) {
      XF_HIP(hipStreamSynchronize(c->stream));
      XF_HIP(hipFree(c->ws[slot]));
      c->ws[slot] = nullptr;
      c->ws_cap[slot] = 0;
    }
    XF_HIP(hipMalloc(&c->ws[slot], bytes));
    c->ws_cap[slot] = bytes;
  }
  return c->ws[slot];
}

const std::vector<uint8_t> &host_codes(xf_grid g) {
  if (g->codes_h.empty()) {
    g->codes_h.resize((size_t)g->ncell);
    XF_HIP(hipMemcpy(g->codes_h.data(), g->d_codes, (size_t)g->ncell, hipMemcpyDeviceToHost));
  }
  return g->codes_h;
}

// A device copy of the host steps h[0..cnt) for the grid's slot, uploaded
// only when they differ from the last ones passed.
const double *device_steps(xf_ctx c, xf_grid g, int slot, const double *h, int cnt) {
  std::vector<double> &hs = g->steps_h[slot];
  if ((int)hs.size() == cnt && g->steps_d[slot] && std::equal(hs.begin(), hs.end(), h)) return g->steps_d[slot];
  if ((int)hs.size() != cnt || !g->steps_d[slot]) {
    if (g->steps_d[slot]) {
      XF_HIP(hipStreamSynchronize(c->stream));
      XF_HIP(hipFree(g->steps_d[slot]));
      g->steps_d[slot] = nullptr;
    }
    XF_HIP(hipMalloc(&g->steps_d[slot], sizeof(double) * (cnt > 0 ? cnt : 1)));
  }
  hs.assign(h, h + cnt);
  // from the cached host vector, which outlives the stream-ordered copy
  XF_HIP(hipMemcpyAsync(g->steps_d[slot], hs.data(), sizeof(double) * cnt, hipMemcpyHostToDevice, c->stream));
  return g->steps_d[slot];
}

const char *err_bit_message(int bits) {
  if (bits & EB_VC_DEFINED) return "TVPProblem::SolveBySplitting: Defined border condition checking error.";
  if (bits & EB_VC_PREDEF)
    return "TVPProblem::SolveBySplitting: PreDefined border condition checking error - value incorrect.";
  return "unknown device error";
}

// Reads and clears the context's error word (the stream is synchronised).
static int take_errors(xf_ctx c) {
  int e = 0;
  XF_HIP(hipMemcpyAsync(&e, c->d_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  XF_HIP(hipStreamSynchronize(c->stream));
  if (e) XF_HIP(hipMemsetAsync(c->d_err, 0, sizeof(int), c->stream));
  return e;
}

extern "C" int xf_set_deferred_errors(xf_ctx c, int on) {
  if (!c) return XF_EINVAL;
  c->defer_errors = on != 0;
  return XF_OK;
}

extern "C" int xf_check_errors(xf_ctx c) {
  return guarded(c, [&] {
    const int e = take_errors(c);
    XF_REQUIRE(e == 0, XF_EINVAL, err_bit_message(e));
  });
}

void raise_device_errors(xf_ctx c) {
  if (c->defer_errors) return;
  const int e = take_errors(c);
  XF_REQUIRE(e == 0, XF_EINVAL, err_bit_message(e));
}

extern "C" int xf_synchronize(xf_ctx c) {
  return guarded(c, [&] { XF_HIP(hipStreamSynchronize(c->stream)); });
}

extern "C" int xf_malloc(xf_ctx c, size_t bytes, void **p) {
  return guarded(c, [&] {
    XF_REQUIRE(p, XF_EINVAL, "xf_malloc: null out pointer");
    hipError_t e = hipMalloc(p, bytes ? bytes : 1);
    XF_REQUIRE(e == hipSuccess, XF_ENOMEM, "xf_malloc: out of device memory");
  });
}

extern "C" int xf_free(xf_ctx c, void *p) {
  return guarded(c, [&] { XF_HIP(hipFree(p)); });
}

extern "C" int xf_memcpy(xf_ctx c, void *dst, const void *src, size_t bytes, int kind) {
  return guarded(c, [&] {
    hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost
                                                                     : hipMemcpyDeviceToDevice;
    XF_REQUIRE(kind >= 0 && kind <= 2, XF_EINVAL, "xf_memcpy: bad kind");
    XF_HIP(hipMemcpyAsync(dst, src, bytes, k, c->stream));
    XF_HIP(hipStreamSynchronize(c->stream));
  });
}

// ------------------------------------------------------------------- grid
// Space mask (XProblem.cpp:4155-4182), flag byte and row validity for one cell.
// Validity mirrors the reference's throws: Actual rows read all six
// neighbours; PreDefined rows need an axis-aligned unit normal
// (XOperator.cpp:1036-1055); PreNormal rows need a non-zero normal
// (XOperator.cpp:1074-1080); Normal and PreEquation rows are unknowns that
// the operator rejects (XOperator.cpp:1120-1123).
__global__ void k_grid_flags(int n, int l, int m, const uint8_t *__restrict__ codes, const double *__restrict__ nrm,
                             const uint8_t *__restrict__ hasn, uint8_t *__restrict__ space,
                             uint8_t *__restrict__ flags, unsigned long long *__restrict__ counts) {
  int64_t nc = (int64_t)n * l * m;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x) {
    int k = (int)(c % m);
    int j = (int)((c / m) % l);
    int i = (int)(c / ((int64_t)l * m));
    uint8_t code = codes[c];
    uint8_t sp = (code == PT_ACTUAL || code == PT_PREDEFINED || code == PT_NORMAL || code == PT_PRENORMAL ||
                  code == PT_PREEQUATION) ? 1 : 0;
    space[c] = sp;
    uint8_t f = 0;
    int64_t sI = (int64_t)l * m, sJ = m;
    if (i > 0 && codes[c - sI] != PT_FICTIVE) f |= FL_IM;
    if (i < n - 1 && codes[c + sI] != PT_FICTIVE) f |= FL_IP;
    if (j > 0 && codes[c - sJ] != PT_FICTIVE) f |= FL_JM;
    if (j < l - 1 && codes[c + sJ] != PT_FICTIVE) f |= FL_JP;
    if (k > 0 && codes[c - 1] != PT_FICTIVE) f |= FL_KM;
    if (k < m - 1 && codes[c + 1] != PT_FICTIVE) f |= FL_KP;
    if (code == PT_ACTUAL) f |= FL_ACTUAL;
    if (sp) f |= FL_UNKNOWN;
    flags[c] = f;
    if (!sp) continue;
    bool interior = i >= 1 && j >= 1 && k >= 1 && i <= n - 2 && j <= l - 2 && k <= m - 2;
    bool bad_lap = false, bad_div = false;
    if (code == PT_ACTUAL) {
      bad_lap = bad_div = !interior;
    } else if (code == PT_PREDEFINED || code == PT_PRENORMAL) {
      atomicAdd(&counts[2], 1ull);
      if (!nrm || !hasn || !hasn[c]) {
        bad_lap = bad_div = true;
      } else {
        double nx = nrm[3 * c], ny = nrm[3 * c + 1], nz = nrm[3 * c + 2];
        if (nx == 0 && ny == 0 && nz == 0) bad_lap = true;
        if (code == PT_PREDEFINED) {
          if (nx != 0) { if (ny != 0 || nz != 0 || fabs(nx) != 1) bad_lap = true; }
          else if (ny != 0) { if (nx != 0 || nz != 0 || fabs(ny) != 1) bad_lap = true; }
          else if (nz != 0) { if (nx != 0 || ny != 0 || fabs(nz) != 1) bad_lap = true; }
          // neighbour read along the normal
          if ((nx > 0 && i < 1) || (nx < 0 && i > n - 2) || (ny > 0 && j < 1) || (ny < 0 && j > l - 2) ||
              (nz > 0 && k < 1) || (nz < 0 && k > m - 2))
            bad_lap = true;
        } else {
          // face averages read i + {-1,0,1} where the normal component is
          // non-zero and i + {0,1} in the tangential offsets
          if ((nx > 0 && i < 1) || (nx < 0 && i > n - 2) || (ny > 0 && j < 1) || (ny < 0 && j > l - 2) ||
              (nz > 0 && k < 1) || (nz < 0 && k > m - 2))
            bad_lap = true;
        }
        bad_div = bad_lap || (code == PT_PRENORMAL && !interior);
      }
    } else {
      atomicAdd(&counts[2], 1ull);
      bad_lap = bad_div = true;  // Normal / PreEquation rows: operator throws
    }
    if (bad_lap) atomicAdd(&counts[0], 1ull);
    if (bad_div) atomicAdd(&counts[1], 1ull);
  }
}

__global__ void k_rn_finish(int64_t nc, const uint8_t *__restrict__ space, int32_t *__restrict__ rn) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x)
    if (!space[c]) rn[c] = 0;
}

struct ToInt {
  __host__ __device__ int32_t operator()(uint8_t v) const { return (int32_t)v; }
};

static void coeffs(const std::vector<double> &h, int cnt, std::vector<double> &cd, std::vector<float> &cf) {
  cd.assign(2 * cnt, 0.0);
  cf.assign(2 * cnt, 0.0f);
  for (int i = 1; i + 1 < cnt; i++) {
    double hh = (h[i - 1] + h[i]) / 2;
    cd[2 * i] = 1.0 / (h[i - 1] * hh);
    cd[2 * i + 1] = 1.0 / (h[i] * hh);
    cf[2 * i] = (float)cd[2 * i];
    cf[2 * i + 1] = (float)cd[2 * i + 1];
  }
}

template <class T>
static T *upload(const std::vector<T> &v) {
  T *d = nullptr;
  XF_HIP(hipMalloc(&d, sizeof(T) * (v.size() ? v.size() : 1)));
  if (!v.empty()) XF_HIP(hipMemcpy(d, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
  return d;
}

void free_grid(xf_grid g) {
  if (!g) return;
  free_mg(g->mg);
  hipFree(g->d_hx); hipFree(g->d_hy); hipFree(g->d_hz);
  hipFree(g->d_cx); hipFree(g->d_cy); hipFree(g->d_cz);
  hipFree(g->d_cxf); hipFree(g->d_cyf); hipFree(g->d_czf);
  hipFree(g->d_codes); hipFree(g->d_flags); hipFree(g->d_space); hipFree(g->d_rn);
  hipFree(g->d_normals); hipFree(g->d_has_normal); hipFree(g->d_rho);
  hipFree(g->d_codes_kt);
  hipFree(g->d_lod_border);
  hipFree(g->d_predef);
  hipFree(g->d_other_rows);
  for (int k = 0; k < 4; k++) hipFree(g->steps_d[k]);
  free_cg_work(g->work);
  delete g;
}

// Where TVPProblem::SolveBySplitting would throw while building the
// transport terms of each equation on a grid with these point types
// (XProblem.cpp:5096-5300 velocity components, 5896-5940 concentration), and
// Actual points whose stencil would leave the array.
static void transport_checks(xf_grid g, const int32_t *codes, const double *normals, const int32_t *has_normal) {
  const int n = g->n, l = g->l, m = g->m;
  const int E[3] = {n, l, m};
  const char *vel_type = "Error in TVPProblem::SolveBySplitting: Method is not implemented for this point border "
                         "type of velocity.";
  const char *vel_eq = "Error in TVPProblem::SolveBySplitting: Method is not implemented for this point border type "
                       "of velocity and this equation number.";
  const char *vel_normal = "Error in TVPProblem::SolveBySplitting: Method is not implemented for this point border "
                           "type of velocity and this equation number and this normal vector.";
  const char *conc = "Error in TVPProblem::SolveBySplitting: Unknown concentration point mask type.";
  const char *edge = "xf_transport_velocity: an Actual point on the array border (its averages leave the array)";
  for (int ax = 0; ax < 4; ax++) {
    g->tv_err[ax].clear();
    g->tv_visterm_ok[ax] = ax > 0;
  }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < l; j++)
      for (int k = 0; k < m; k++) {
        const int64_t c = ((int64_t)i * l + j) * m + k;
        const int code = codes[c];
        const int idx[3] = {i, j, k};
        auto fail = [&](int ax, const char *msg) {
          if (g->tv_err[ax].empty()) g->tv_err[ax] = msg;
        };
        // concentration
        if (code == PT_ACTUAL) {
          if (i < 1 || j < 1 || k < 1) fail(0, edge);
        } else if (!(code == PT_FICTIVE || code == PT_PREDEFINED || code == PT_PRENORMAL)) {
          fail(0, conc);
        }
        // velocity components
        for (int a = 0; a < 3; a++) {
          const int ax = a + 1;
          if (code == PT_ACTUAL) {
            bool ok = idx[a] + 1 < E[a];
            for (int e = 0; e < 3; e++)
              if (e != a && idx[e] < 1) ok = false;
            if (!ok) fail(ax, edge);
            if (i < 1 || j < 1 || k < 1 || i + 1 >= n || j + 1 >= l || k + 1 >= m) g->tv_visterm_ok[ax] = false;
          } else if (code == PT_NORMAL) {
            if (ax != 1) {
              fail(ax, vel_eq);
            } else {
              const bool have = normals && (!has_normal || has_normal[c]);
              if (!have || normals[3 * c] == 0 || normals[3 * c + 1] != 0 || normals[3 * c + 2] != 0)
                fail(ax, vel_normal);
            }
          } else if (code == PT_EQUATION || code == PT_PRENORMAL || code == PT_PREEQUATION) {
            fail(ax, vel_type);
          }
        }
      }
}

extern "C" int xf_grid_create(xf_ctx c, int n, int l, int m, const double *hx, const double *hy,
                              const double *hz, const int32_t *codes, const double *normals,
                              const int32_t *has_normal, xf_grid *out) {
  xf_grid g = nullptr;
  int st = guarded(c, [&] {
    XF_REQUIRE(c && out && hx && hy && hz && codes, XF_EINVAL, "xf_grid_create: null argument");
    XF_REQUIRE(n >= 2 && l >= 2 && m >= 2, XF_EINVAL, "xf_grid_create: each axis needs at least 2 nodes");
    int64_t nc = (int64_t)n * l * m;
    XF_REQUIRE(nc < (int64_t)1 << 31, XF_EINVAL, "xf_grid_create: grid too large for 32-bit Rn indices");
    XF_HIP(hipSetDevice(c->device));
    g = new xf_grid_s();
    g->n = n; g->l = l; g->m = m; g->ncell = nc;
    g->hx.assign(hx, hx + n - 1);
    g->hy.assign(hy, hy + l - 1);
    g->hz.assign(hz, hz + m - 1);
    std::vector<uint8_t> cb(nc);
    for (int64_t q = 0; q < nc; q++) {
      XF_REQUIRE(codes[q] >= 0 && codes[q] <= 8 && codes[q] != PT_BORDER, XF_EINVAL,
                 "xf_grid_create: unknown point type in mask");
      cb[q] = (uint8_t)codes[q];
    }
    transport_checks(g, codes, normals, has_normal);
    g->d_hx = upload(g->hx); g->d_hy = upload(g->hy); g->d_hz = upload(g->hz);
    std::vector<double> cd;
    std::vector<float> cf;
    coeffs(g->hx, n, cd, cf); g->d_cx = upload(cd); g->d_cxf = upload(cf);
    coeffs(g->hy, l, cd, cf); g->d_cy = upload(cd); g->d_cyf = upload(cf);
    coeffs(g->hz, m, cd, cf); g->d_cz = upload(cd); g->d_czf = upload(cf);
    g->d_codes = upload(cb);
    if (normals) {
      XF_HIP(hipMalloc(&g->d_normals, sizeof(double) * nc * 3));
      XF_HIP(hipMemcpy(g->d_normals, normals, sizeof(double) * nc * 3, hipMemcpyHostToDevice));
      std::vector<uint8_t> hn(nc, 1);
      if (has_normal)
        for (int64_t q = 0; q < nc; q++) hn[q] = has_normal[q] ? 1 : 0;
      g->d_has_normal = upload(hn);
    }
    XF_HIP(hipMalloc(&g->d_flags, nc));
    XF_HIP(hipMalloc(&g->d_space, nc));
    XF_HIP(hipMalloc(&g->d_rn, sizeof(int32_t) * nc));
    unsigned long long *d_counts = nullptr;
    XF_HIP(hipMalloc(&d_counts, sizeof(unsigned long long) * 3));
    XF_HIP(hipMemsetAsync(d_counts, 0, sizeof(unsigned long long) * 3, c->stream));
    int nb = grid_blocks(nc);
    if (nb > 65536) nb = 65536;
    hipLaunchKernelGGL(k_grid_flags, dim3(nb), dim3(kBlock), 0, c->stream, n, l, m, g->d_codes, g->d_normals,
                       g->d_has_normal, g->d_space, g->d_flags, d_counts);
    check_kernel("k_grid_flags");
    // Rn numbering: inclusive prefix sum of the space mask, zeroed off the mask.
    hipcub::TransformInputIterator<int32_t, ToInt, const uint8_t *> it(g->d_space, ToInt());
    size_t tmp_bytes = 0;
    XF_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tmp_bytes, it, g->d_rn, (int)nc, c->stream));
    void *tmp = nullptr;
    XF_HIP(hipMalloc(&tmp, tmp_bytes));
    XF_HIP(hipcub::DeviceScan::InclusiveSum(tmp, tmp_bytes, it, g->d_rn, (int)nc, c->stream));
    hipLaunchKernelGGL(k_rn_finish, dim3(nb), dim3(kBlock), 0, c->stream, nc, g->d_space, g->d_rn);
    check_kernel("k_rn_finish");
    unsigned long long counts[3];
    int32_t last = 0;
    XF_HIP(hipMemcpyAsync(counts, d_counts, sizeof(counts), hipMemcpyDeviceToHost, c->stream));
    XF_HIP(hipStreamSynchronize(c->stream));
    hipFree(tmp);
    hipFree(d_counts);
    // n_unknowns = the largest Rn index
    int32_t *d_max = nullptr;
    size_t mb = 0;
    XF_HIP(hipMalloc(&d_max, sizeof(int32_t)));
    XF_HIP(hipcub::DeviceReduce::Max(nullptr, mb, g->d_rn, d_max, (int)nc, c->stream));
    void *mt = nullptr;
    XF_HIP(hipMalloc(&mt, mb));
    XF_HIP(hipcub::DeviceReduce::Max(mt, mb, g->d_rn, d_max, (int)nc, c->stream));
    XF_HIP(hipMemcpyAsync(&last, d_max, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    XF_HIP(hipStreamSynchronize(c->stream));
    hipFree(mt);
    hipFree(d_max);
    g->nunk = last;
    g->bad_rows_lap = (int64_t)counts[0];
    g->bad_rows_div = (int64_t)counts[1];
    g->special_rows = (int64_t)counts[2];
  });
  if (st != XF_OK) {
    free_grid(g);
    return st;
  }
  *out = g;
  return XF_OK;
}

extern "C" int xf_grid_destroy(xf_ctx c, xf_grid g) {
  return guarded(c, [&] {
    XF_HIP(hipStreamSynchronize(c->stream));
    free_grid(g);
  });
}

extern "C" int xf_grid_info(xf_grid g, int64_t *ncell, int64_t *nunk) {
  if (!g) return XF_EINVAL;
  if (ncell) *ncell = g->ncell;
  if (nunk) *nunk = g->nunk;
  return XF_OK;
}

// Activity table rows (XOperator.cpp:837-936).
__global__ void k_activity(int n, int l, int m, const uint8_t *__restrict__ space, const int32_t *__restrict__ rn,
                           int32_t *__restrict__ act) {
  int64_t nc = (int64_t)n * l * m;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x) {
    if (!space[c]) continue;
    int k = (int)(c % m);
    int j = (int)((c / m) % l);
    int i = (int)(c / ((int64_t)l * m));
    int32_t row[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int cnt = 1;
    row[1] = rn[c];
    const int d[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
    for (int q = 0; q < 6; q++) {
      int a = i + d[q][0], b = j + d[q][1], e = k + d[q][2];
      if (a >= 0 && a < n && b >= 0 && b < l && e >= 0 && e < m) {
        int64_t cc = ((int64_t)a * l + b) * m + e;
        if (space[cc]) row[++cnt] = rn[cc];
      }
    }
    row[0] = cnt;
    int32_t *dst = act + (int64_t)(rn[c] - 1) * 8;
    for (int q = 0; q < 8; q++) dst[q] = row[q];
  }
}

extern "C" int xf_grid_export(xf_ctx c, xf_grid g, int32_t *space_mask, int32_t *rn, int32_t *act) {
  return guarded(c, [&] {
    XF_REQUIRE(g, XF_EINVAL, "xf_grid_export: null grid");
    int64_t nc = g->ncell;
    if (space_mask) {
      std::vector<uint8_t> s(nc);
      XF_HIP(hipMemcpyAsync(s.data(), g->d_space, nc, hipMemcpyDeviceToHost, c->stream));
      XF_HIP(hipStreamSynchronize(c->stream));
      for (int64_t q = 0; q < nc; q++) space_mask[q] = s[q];
    }
    if (rn) {
      XF_HIP(hipMemcpyAsync(rn, g->d_rn, sizeof(int32_t) * nc, hipMemcpyDeviceToHost, c->stream));
      XF_HIP(hipStreamSynchronize(c->stream));
    }
    if (act && g->nunk > 0) {
      int32_t *d_act = nullptr;
      XF_HIP(hipMalloc(&d_act, sizeof(int32_t) * g->nunk * 8));
      int nb = grid_blocks(nc);
      if (nb > 65536) nb = 65536;
      hipLaunchKernelGGL(k_activity, dim3(nb), dim3(kBlock), 0, c->stream, g->n, g->l, g->m, g->d_space, g->d_rn,
                         d_act);
      check_kernel("k_activity");
      XF_HIP(hipMemcpyAsync(act, d_act, sizeof(int32_t) * g->nunk * 8, hipMemcpyDeviceToHost, c->stream));
      XF_HIP(hipStreamSynchronize(c->stream));
      hipFree(d_act);
    }
  });
}

extern "C" int xf_grid_set_density(xf_ctx c, xf_grid g, const double *rho) {
  return guarded(c, [&] {
    XF_REQUIRE(g, XF_EINVAL, "xf_grid_set_density: null grid");
    if (!rho) {
      g->use_rho = false;
      return;
    }
    if (!g->d_rho) XF_HIP(hipMalloc(&g->d_rho, sizeof(double) * g->ncell));
    XF_HIP(hipMemcpyAsync(g->d_rho, rho, sizeof(double) * g->ncell, hipMemcpyDeviceToDevice, c->stream));
    g->use_rho = true;
  });
}

// --------------------------------------------------------------- profiling
int prof_begin(xf_ctx c) {
  // a fresh (start, stop) pair per launch; prof_collect destroys them
  int idx = (int)c->prof_events.size();
  hipEvent_t a, b;
  XF_HIP(hipEventCreate(&a));
  XF_HIP(hipEventCreate(&b));
  c->prof_events.push_back(a);
  c->prof_events.push_back(b);
  XF_HIP(hipEventRecord(a, c->stream));
  return idx;
}

void prof_end(xf_ctx c, int ev) { XF_HIP(hipEventRecord(c->prof_events[ev + 1], c->stream)); }

void prof_collect(xf_ctx c, std::vector<ProfMark> &marks, const std::map<std::string, long long> &executed) {
  std::map<std::string, long long> seen;
  for (auto &mk : marks) {
    long long lim = 0;
    auto it = executed.find(mk.name);
    if (it != executed.end()) lim = it->second;
    if (seen[mk.name]++ < lim) {
      float ms = 0;
      XF_HIP(hipEventElapsedTime(&ms, c->prof_events[mk.ev], c->prof_events[mk.ev + 1]));
      auto &acc = c->prof_acc[mk.name];
      acc.first += ms;
      acc.second += 1;
    }
  }
  for (auto e : c->prof_events) hipEventDestroy(e);
  c->prof_events.clear();
  marks.clear();
}

extern "C" int xf_profile_enable(xf_ctx c, int on) {
  if (!c) return XF_EINVAL;
  c->prof = on != 0;
  return XF_OK;
}

extern "C" int xf_profile_reset(xf_ctx c) {
  if (!c) return XF_EINVAL;
  c->prof_acc.clear();
  return XF_OK;
}

extern "C" int xf_profile_read(xf_ctx c, const char *kernel, double *avg_ms, int64_t *launches) {
  if (!c || !kernel) return XF_EINVAL;
  auto it = c->prof_acc.find(kernel);
  double a = 0;
  long long n = 0;
  if (it != c->prof_acc.end()) {
    n = it->second.second;
    a = n ? it->second.first / (double)n : 0;
  }
  if (avg_ms) *avg_ms = a;
  if (launches) *launches = n;
  return XF_OK;
}
