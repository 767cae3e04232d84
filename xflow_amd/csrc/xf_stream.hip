// Streaming-bandwidth ceilings of this GPU for the access patterns of the
// solver kernels: a read-only reduction (the dots kernel reads r and flags),
// a copy (one read, one write) and two reads with two writes in place (the
// CG update reads r and p and writes both).  Grid-stride loops over 16-byte
// vectors, blocks a multiple of the XCD count.  Measurement only.
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
    auto launch = [&]() {
      if (kind == 0) hipLaunchKernelGGL(k_stream_read, dim3(nb), dim3(256), 0, c->stream, n2, a, part);
      else if (kind == 1) hipLaunchKernelGGL(k_stream_copy, dim3(nb), dim3(256), 0, c->stream, n2, a, b);
      else hipLaunchKernelGGL(k_stream_2r2w, dim3(nb), dim3(256), 0, c->stream, n2, a, b, 0.0);
    };
    launch();
    check_kernel("stream ceiling");
    XF_HIP(hipEventRecord(c->ev0, c->stream));
    for (int r = 0; r < reps; r++) launch();
    XF_HIP(hipEventRecord(c->ev1, c->stream));
    XF_HIP(hipEventSynchronize(c->ev1));
    float ms = 0;
    XF_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    const double moved = (kind == 0 ? 1.0 : (kind == 1 ? 2.0 : 4.0)) * 16.0 * (double)n2;
    *gbs = moved / (ms / reps * 1e-3) / 1e9;
  });
}
