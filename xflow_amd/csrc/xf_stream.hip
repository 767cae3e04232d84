// Streaming-bandwidth ceilings of this GPU for the access patterns of the
// solver kernels: a read-only reduction (the dots kernel reads r and flags),
// a copy (one read, one write) and two reads with two writes in place (the
// CG update reads r and p and writes both).  Each pattern is timed as a
// grid-stride loop over 16-byte vectors (blocks a multiple of the XCD count)
// and as a flat launch of one vector per thread (plain and non-temporal
// accesses); the ceiling is the fastest form.  Measurement only.
#include <algorithm>

#include "xf_internal.h"

namespace {

__global__ __launch_bounds__(256) void k_stream_read(int64_t n2, const double2 *__restrict__ a,
                                                     double *__restrict__ part) {
  __shared__ double lds[8];
  double acc = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n2; i += (int64_t)gridDim.x * blockDim.x) {
    const double2 v = a[i];
    acc += v.x + v.y;
  }
  acc = block_sum(acc, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

__global__ __launch_bounds__(256) void k_stream_copy(int64_t n2, const double2 *__restrict__ a,
                                                     double2 *__restrict__ b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n2; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}

__global__ __launch_bounds__(256) void k_stream_2r2w(int64_t n2, double2 *__restrict__ a, double2 *__restrict__ b,
                                                     double s) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n2; i += (int64_t)gridDim.x * blockDim.x) {
    const double2 x = a[i], y = b[i];
    const double2 q = make_double2(x.x + s * y.x, x.y + s * y.y);
    b[i] = q;
    a[i] = make_double2(x.x - s * q.x, x.y - s * q.y);
  }
}

typedef double dv2 __attribute__((ext_vector_type(2)));

// one 16-byte vector per thread; NT: non-temporal loads and stores
template <int NT>
__global__ __launch_bounds__(256) void k_stream_read_flat(int64_t n2, const dv2 *__restrict__ a,
                                                          double *__restrict__ part) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  double acc = 0;
  if (i < n2) {
    const dv2 v = NT ? __builtin_nontemporal_load(a + i) : a[i];
    acc = v.x + v.y;
  }
  // a value no run produces (the buffers hold zeros) keeps the loads alive
  if (acc == 1.5) part[0] = acc;
}

template <int NT>
__global__ __launch_bounds__(256) void k_stream_copy_flat(int64_t n2, const dv2 *__restrict__ a,
                                                          dv2 *__restrict__ b) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n2) return;
  if (NT) {
    __builtin_nontemporal_store(__builtin_nontemporal_load(a + i), b + i);
  } else {
    b[i] = a[i];
  }
}

template <int NT>
__global__ __launch_bounds__(256) void k_stream_2r2w_flat(int64_t n2, dv2 *__restrict__ a, dv2 *__restrict__ b,
                                                          double s) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n2) return;
  const dv2 x = NT ? __builtin_nontemporal_load(a + i) : a[i];
  const dv2 y = NT ? __builtin_nontemporal_load(b + i) : b[i];
  const dv2 q = x + s * y;
  const dv2 r = x - s * q;
  if (NT) {
    __builtin_nontemporal_store(q, b + i);
    __builtin_nontemporal_store(r, a + i);
  } else {
    b[i] = q;
    a[i] = r;
  }
}

}  // namespace

extern "C" int xf_stream_ceiling(xf_ctx c, int kind, size_t bytes, int reps, double *gbs) {
  return guarded(c, [&] {
    XF_REQUIRE(gbs && kind >= 0 && kind <= 2 && reps > 0 && bytes >= 4096, XF_EINVAL,
               "xf_stream_ceiling: kind 0..2, reps > 0, bytes >= 4096");
    const int64_t n2 = (int64_t)(bytes / 16);
    double2 *a = nullptr, *b = nullptr;
    double *part = nullptr;
    const int nb = 65536;
    XF_HIP(hipMalloc(&a, n2 * 16));
    struct Free {
      void *p[3];
      ~Free() {
        for (void *q : p)
          if (q) hipFree(q);
      }
    } fr{{a, nullptr, nullptr}};
    XF_HIP(hipMalloc(&b, n2 * 16));
    fr.p[1] = b;
    XF_HIP(hipMalloc(&part, nb * sizeof(double)));
    fr.p[2] = part;
    XF_HIP(hipMemsetAsync(a, 0, n2 * 16, c->stream));
    XF_HIP(hipMemsetAsync(b, 0, n2 * 16, c->stream));
    const int64_t nflat = (n2 + 255) / 256;
    XF_REQUIRE(nflat < ((int64_t)1 << 31), XF_EINVAL, "xf_stream_ceiling: buffer too large for a flat launch");
    dv2 *fa = (dv2 *)a, *fb = (dv2 *)b;
    auto launch = [&](int form) {
      const dim3 g = form == 0 ? dim3(nb) : dim3((unsigned)nflat);
      if (kind == 0) {
        if (form == 0) hipLaunchKernelGGL(k_stream_read, g, dim3(256), 0, c->stream, n2, a, part);
        else if (form == 1) hipLaunchKernelGGL(k_stream_read_flat<0>, g, dim3(256), 0, c->stream, n2, fa, part);
        else hipLaunchKernelGGL(k_stream_read_flat<1>, g, dim3(256), 0, c->stream, n2, fa, part);
      } else if (kind == 1) {
        if (form == 0) hipLaunchKernelGGL(k_stream_copy, g, dim3(256), 0, c->stream, n2, a, b);
        else if (form == 1) hipLaunchKernelGGL(k_stream_copy_flat<0>, g, dim3(256), 0, c->stream, n2, fa, fb);
        else hipLaunchKernelGGL(k_stream_copy_flat<1>, g, dim3(256), 0, c->stream, n2, fa, fb);
      } else {
        if (form == 0) hipLaunchKernelGGL(k_stream_2r2w, g, dim3(256), 0, c->stream, n2, a, b, 0.0);
        else if (form == 1) hipLaunchKernelGGL(k_stream_2r2w_flat<0>, g, dim3(256), 0, c->stream, n2, fa, fb, 0.0);
        else hipLaunchKernelGGL(k_stream_2r2w_flat<1>, g, dim3(256), 0, c->stream, n2, fa, fb, 0.0);
      }
    };
    const double moved = (kind == 0 ? 1.0 : (kind == 1 ? 2.0 : 4.0)) * 16.0 * (double)n2;
    double best = 0;
    for (int form = 0; form < 3; form++) {
      launch(form);
      check_kernel("stream ceiling");
      XF_HIP(hipEventRecord(c->ev0, c->stream));
      for (int r = 0; r < reps; r++) launch(form);
      XF_HIP(hipEventRecord(c->ev1, c->stream));
      XF_HIP(hipEventSynchronize(c->ev1));
      float ms = 0;
      XF_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
      best = std::max(best, moved / (ms / reps * 1e-3) / 1e9);
    }
    *gbs = best;
  });
}
