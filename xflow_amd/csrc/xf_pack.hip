// T1dPackOperator on the device (XOperator.cpp:1135-1571): the operator's
// matrix in packed rows (CSR) over the Rn unknowns, for A and for A^T, and
// the products with Rn-ordered vectors (XOperator.cpp:1551-1568).
//
// Row r of A lists, in the order of activity-table row r (the row's own
// unknown, then the i-1, i+1, j-1, j+1, k-1, k+1 neighbours that are
// unknowns), the non-zero coefficients A[r][q]; row r of A^T lists, in the same
// order over q, the non-zero A[q][r].  A coefficient is the reference row
// evaluated on a unit vector (the reference probes columns the same way,
// XOperator.cpp:1207-1224), so the packed values are the reference's bits.
// One thread per row: count, exclusive scan, fill; the product sums each row
// in stored order like the reference.
#include <hipcub/hipcub.hpp>

#include "xf_rows.h"

namespace {

struct UnitF {
  int64_t at;
  __device__ double operator()(int64_t q) const { return q == at ? 1.0 : 0.0; }
};

__global__ void k_pack_cells(int64_t nc, const int32_t *__restrict__ rn, int64_t *__restrict__ cell) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x)
    if (rn[c] > 0) cell[rn[c] - 1] = c;
}

// FILL = false: counts[r] = number of non-zeros of row r; true: write them at ptr[r]
template <bool FILL>
__global__ void k_pack_rows(GridView g, const int64_t *__restrict__ cell, int64_t nunk, int transpose,
                            int32_t *__restrict__ counts, const int32_t *__restrict__ ptr, int32_t *__restrict__ col,
                            double *__restrict__ val) {
  const int64_t sI = (int64_t)g.l * g.m, sJ = g.m;
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < nunk; r += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = cell[r];
    int i, j, k;
    decompose(c, g.l, g.m, i, j, k);
    const int di[7] = {0, -1, 1, 0, 0, 0, 0}, dj[7] = {0, 0, 0, -1, 1, 0, 0}, dk[7] = {0, 0, 0, 0, 0, -1, 1};
    int cnt = 0;
    int64_t q = FILL ? ptr[r] : 0;
    for (int e = 0; e < 7; e++) {
      const int a = i + di[e], b = j + dj[e], w = k + dk[e];
      if (a < 0 || a >= g.n || b < 0 || b >= g.l || w < 0 || w >= g.m) continue;
      const int64_t cc = c + di[e] * sI + dj[e] * sJ + dk[e];
      if (!g.space[cc]) continue;
      const double v = transpose ? row_value_t(g, UnitF{c}, cc, a, b, w) : row_value_t(g, UnitF{cc}, c, i, j, k);
      if (v != 0) {
        if (FILL) {
          val[q] = v;
          col[q] = g.rn[cc] - 1;
          q++;
        }
        cnt++;
      }
    }
    if (!FILL) counts[r] = cnt;
  }
}

// y = A x (+ b or - b with sign_b); st (may be null): return once a solver is done
__global__ void k_pack_apply(int64_t nunk, const int32_t *__restrict__ ptr, const int32_t *__restrict__ col,
                             const double *__restrict__ val, const double *__restrict__ x,
                             const double *__restrict__ b, double sign_b, double *__restrict__ y,
                             const DevState *st) {
  if (st && st->done) return;
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < nunk; r += (int64_t)gridDim.x * blockDim.x) {
    double res = 0;
    for (int32_t q = ptr[r]; q < ptr[r + 1]; q++) res = res + val[q] * x[col[q]];
    if (sign_b > 0) res = res + b[r];
    else if (sign_b < 0) res = res - b[r];
    y[r] = res;
  }
}

// (A x, w) into *out
__global__ void k_pack_apply_dot(int64_t nunk, const int32_t *__restrict__ ptr, const int32_t *__restrict__ col,
                                 const double *__restrict__ val, const double *__restrict__ x,
                                 const double *__restrict__ w, double *partials, unsigned *counter, double *out,
                                 const DevState *st) {
  __shared__ double lds[16];
  if (st->done) return;
  double acc = 0;
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < nunk; r += (int64_t)gridDim.x * blockDim.x) {
    double res = 0;
    for (int32_t q = ptr[r]; q < ptr[r + 1]; q++) res = res + val[q] * x[col[q]];
    acc += res * w[r];
  }
  double bs = block_sum(acc, lds);
  double tot;
  if (grid_sum_last(bs, partials, counter, &tot, lds) && threadIdx.x == 0) *out = tot;
}

__global__ void k_gather_rn(int64_t nunk, const int64_t *__restrict__ cell, const double *__restrict__ f,
                            double *__restrict__ v) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < nunk; r += (int64_t)gridDim.x * blockDim.x)
    v[r] = f[cell[r]];
}

__global__ void k_scatter_rn(int64_t nunk, const int64_t *__restrict__ cell, const double *__restrict__ v,
                             double *__restrict__ f) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < nunk; r += (int64_t)gridDim.x * blockDim.x)
    f[cell[r]] = v[r];
}

int pack_blocks(int64_t n) {
  int64_t b = (n + kBlock - 1) / kBlock;
  return (int)(b > 16384 ? 16384 : (b < 1 ? 1 : b));
}

void free_pack(xf_pack_s *p) {
  if (!p) return;
  hipFree(p->cell);
  for (int t = 0; t < 2; t++) {
    hipFree(p->ptr[t]);
    hipFree(p->col[t]);
    hipFree(p->val[t]);
  }
  delete p;
}

}  // namespace

void launch_pack_apply(xf_ctx c, xf_pack_s *p, int transpose, const double *x, const double *b, double sign_b,
                       double *y, bool gated) {
  const int t = transpose ? 1 : 0;
  hipLaunchKernelGGL(k_pack_apply, dim3(pack_blocks(p->nunk)), dim3(kBlock), 0, c->stream, p->nunk, p->ptr[t],
                     p->col[t], p->val[t], x, b, sign_b, y, gated ? c->state : (const DevState *)nullptr);
  check_kernel("k_pack_apply");
}

void launch_pack_apply_dot(xf_ctx c, xf_pack_s *p, int t, const double *x, const double *w, double *out) {
  int nb = pack_blocks(p->nunk);
  if (nb > 2048) nb = 2048;
  hipLaunchKernelGGL(k_pack_apply_dot, dim3(nb), dim3(kBlock), 0, c->stream, p->nunk, p->ptr[t], p->col[t],
                     p->val[t], x, w, c->partials, c->counters, out, c->state);
  check_kernel("k_pack_apply_dot");
}

extern "C" int xf_pack_create(xf_ctx c, xf_grid g, xf_pack *out) {
  xf_pack_s *p = nullptr;
  int st = guarded(c, [&] {
    XF_REQUIRE(c && g && out, XF_EINVAL, "xf_pack_create: null argument");
    XF_REQUIRE((g->use_rho ? g->bad_rows_div : g->bad_rows_lap) == 0, XF_EINVAL,
               "T1dPackOperator: a row of this mask throws in the reference");
    p = new xf_pack_s();
    p->g = g;
    p->nunk = g->nunk;
    const int64_t n = p->nunk;
    XF_REQUIRE(n > 0, XF_EINVAL, "T1dPackOperator: no unknowns");
    XF_REQUIRE(7 * n < ((int64_t)1 << 31), XF_EINVAL, "T1dPackOperator: too many entries for 32-bit row starts");
    XF_HIP(hipMalloc(&p->cell, sizeof(int64_t) * n));
    hipLaunchKernelGGL(k_pack_cells, dim3(pack_blocks(g->ncell)), dim3(kBlock), 0, c->stream, g->ncell, g->d_rn,
                       p->cell);
    const GridView v = view(g);
    int32_t *counts = nullptr;
    XF_HIP(hipMalloc(&counts, sizeof(int32_t) * (n + 1)));
    for (int t = 0; t < 2; t++) {
      XF_HIP(hipMalloc(&p->ptr[t], sizeof(int32_t) * (n + 1)));
      XF_HIP(hipMemsetAsync(counts, 0, sizeof(int32_t) * (n + 1), c->stream));
      hipLaunchKernelGGL(k_pack_rows<false>, dim3(pack_blocks(n)), dim3(kBlock), 0, c->stream, v, p->cell, n, t,
                         counts, (const int32_t *)nullptr, (int32_t *)nullptr, (double *)nullptr);
      size_t tmp_bytes = 0;
      XF_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, counts, p->ptr[t], (int)(n + 1), c->stream));
      void *tmp = nullptr;
      XF_HIP(hipMalloc(&tmp, tmp_bytes));
      XF_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, counts, p->ptr[t], (int)(n + 1), c->stream));
      int32_t nnz = 0;
      XF_HIP(hipMemcpyAsync(&nnz, p->ptr[t] + n, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
      XF_HIP(hipStreamSynchronize(c->stream));
      hipFree(tmp);
      p->nnz[t] = nnz;
      XF_HIP(hipMalloc(&p->col[t], sizeof(int32_t) * (nnz ? nnz : 1)));
      XF_HIP(hipMalloc(&p->val[t], sizeof(double) * (nnz ? nnz : 1)));
      hipLaunchKernelGGL(k_pack_rows<true>, dim3(pack_blocks(n)), dim3(kBlock), 0, c->stream, v, p->cell, n, t,
                         (int32_t *)nullptr, p->ptr[t], p->col[t], p->val[t]);
      check_kernel("k_pack_rows");
    }
    // an empty row throws in the reference (XOperator.cpp:1239, 1335)
    std::vector<int32_t> hp(n + 1);
    for (int t = 0; t < 2; t++) {
      XF_HIP(hipMemcpy(hp.data(), p->ptr[t], sizeof(int32_t) * (n + 1), hipMemcpyDeviceToHost));
      for (int64_t r = 0; r < n; r++)
        XF_REQUIRE(hp[r + 1] > hp[r], XF_EINVAL, "T1dPackOperator: zero row while constructing the matrix");
    }
    hipFree(counts);
  });
  if (st != XF_OK) {
    free_pack(p);
    return st;
  }
  *out = p;
  return XF_OK;
}

extern "C" int xf_pack_destroy(xf_ctx c, xf_pack p) {
  return guarded(c, [&] {
    if (!p) return;
    XF_HIP(hipStreamSynchronize(c->stream));
    free_pack(p);
  });
}

extern "C" int xf_pack_info(xf_pack p, int64_t *n_unknowns, int64_t *nnz, int64_t *nnz_transpose) {
  if (!p) return XF_EINVAL;
  if (n_unknowns) *n_unknowns = p->nunk;
  if (nnz) *nnz = p->nnz[0];
  if (nnz_transpose) *nnz_transpose = p->nnz[1];
  return XF_OK;
}

extern "C" int xf_pack_export(xf_ctx c, xf_pack p, int transpose, int32_t *counts, double *vals, int32_t *cols) {
  return guarded(c, [&] {
    XF_REQUIRE(p, XF_EINVAL, "xf_pack_export: null pack");
    const int t = transpose ? 1 : 0;
    std::vector<int32_t> hp(p->nunk + 1);
    XF_HIP(hipMemcpy(hp.data(), p->ptr[t], sizeof(int32_t) * (p->nunk + 1), hipMemcpyDeviceToHost));
    if (counts)
      for (int64_t r = 0; r < p->nunk; r++) counts[r] = hp[r + 1] - hp[r];
    if (vals) XF_HIP(hipMemcpy(vals, p->val[t], sizeof(double) * p->nnz[t], hipMemcpyDeviceToHost));
    if (cols) XF_HIP(hipMemcpy(cols, p->col[t], sizeof(int32_t) * p->nnz[t], hipMemcpyDeviceToHost));
  });
}

extern "C" int xf_pack_apply(xf_ctx c, xf_pack p, int transpose, const double *x, double *y) {
  return guarded(c, [&] {
    XF_REQUIRE(p && x && y && x != y, XF_EINVAL, "xf_pack_apply: bad argument");
    launch_pack_apply(c, p, transpose, x, nullptr, 0, y, false);
  });
}

extern "C" int xf_pack_gather(xf_ctx c, xf_pack p, const double *field, double *vec) {
  return guarded(c, [&] {
    XF_REQUIRE(p && field && vec, XF_EINVAL, "xf_pack_gather: null argument");
    hipLaunchKernelGGL(k_gather_rn, dim3(pack_blocks(p->nunk)), dim3(kBlock), 0, c->stream, p->nunk, p->cell, field,
                       vec);
    check_kernel("k_gather_rn");
  });
}

extern "C" int xf_pack_scatter(xf_ctx c, xf_pack p, const double *vec, double *field) {
  return guarded(c, [&] {
    XF_REQUIRE(p && field && vec, XF_EINVAL, "xf_pack_scatter: null argument");
    hipLaunchKernelGGL(k_scatter_rn, dim3(pack_blocks(p->nunk)), dim3(kBlock), 0, c->stream, p->nunk, p->cell, vec,
                       field);
    check_kernel("k_scatter_rn");
  });
}
