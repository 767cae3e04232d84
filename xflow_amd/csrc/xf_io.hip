// On-disk formats of the reference's output path:
//   * the node-centred ASCII VTK structured grid that TVPProblem::SolveBySplitting
//     writes (XProblem.cpp:6132-6379): velocity as the mean of the four faces
//     around a node, pressure / density / concentration / viscosity as the sum of
//     the eight cells around it over the count of non-Fictive ones among them;
//   * the immersed-boundary VTK of TImmersedBoundaryProblem::OutputBoundary
//     (XProblem.cpp:9543-9635);
//   * cnpy's .npy (cnpy.h:66-125, 207-236).
// Node averages run on the GPU a chunk of nodes at a time; the text is
// formatted on the host with the reference's printf formats by several
// threads, each into its own buffer, and the buffers are written in order.
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "xf_internal.h"

namespace {

constexpr int64_t kVtkChunk = (int64_t)1 << 22;  // nodes per device pass

// quantity of one pass: 0 velocity (3 values), 1 pressure, 2 density,
// 3 concentration, 4 viscosity
__global__ void k_vtk_nodes(int n, int l, int m, int64_t q0, int64_t cnt, int what, const uint8_t *__restrict__ mask,
                            const double *__restrict__ a, const double *__restrict__ b,
                            const double *__restrict__ c, double *__restrict__ out) {
  const int NI = n - 2, NJ = l - 2;
  const int64_t sI = (int64_t)l * m, sJ = m;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < cnt; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t q = q0 + t;  // node order of the file: i fastest, then j, then k
    const int i = 1 + (int)(q % NI);
    const int64_t r = q / NI;
    const int j = 1 + (int)(r % NJ), k = 1 + (int)(r / NJ);
    const int64_t x = ((int64_t)i * l + j) * m + k;
    if (what == 0) {
      double uu = 0, vv = 0, ww = 0;
      if (mask[x] != 0) {  // XProblem.cpp:6181-6187
        uu = (a[x] + a[x + sJ] + a[x + 1] + a[x + sJ + 1]) / 4;
        vv = (b[x] + b[x + sI] + b[x + 1] + b[x + sI + 1]) / 4;
        ww = (c[x] + c[x + sI] + c[x + sJ] + c[x + sI + sJ]) / 4;
      }
      out[3 * t] = uu;
      out[3 * t + 1] = vv;
      out[3 * t + 2] = ww;
    } else {
      double v = 0;
      if (mask[x] != 0 && a) {  // XProblem.cpp:6208-6239: the sum takes all eight cells
        const int64_t cell[8] = {x + sI, x + sI + sJ, x + sI + 1, x + sI + sJ + 1, x, x + sJ, x + 1, x + sJ + 1};
        int avr = 0;
        for (int e = 0; e < 8; e++) avr += mask[cell[e]] != 0;
        double s = a[cell[0]];
        for (int e = 1; e < 8; e++) s = s + a[cell[e]];
        v = s / avr;  // avr >= 1: the node's own cell is not Fictive
      }
      out[t] = v;
    }
  }
}

// one(q, buf, cap) formats item q of [0, count); items go to the file in order
template <class F>
static void format_parallel(int64_t count, F one, FILE *fp) {
  const int64_t kBlockItems = 1 << 16;
  unsigned nt = std::thread::hardware_concurrency();
  nt = nt == 0 ? 4 : (nt > 16 ? 16 : nt);
  for (int64_t base = 0; base < count; base += kBlockItems * nt) {
    std::vector<std::string> bufs(nt);
    std::vector<std::thread> th;
    for (unsigned w = 0; w < nt; w++) {
      const int64_t lo = base + (int64_t)w * kBlockItems, hi = std::min(count, lo + kBlockItems);
      if (lo >= hi) break;
      th.emplace_back([&, w, lo, hi] {
        std::string &s = bufs[w];
        s.reserve((size_t)(hi - lo) * 40);
        char tmp[160];
        for (int64_t q = lo; q < hi; q++) {
          const int len = one(q, tmp, sizeof(tmp));
          s.append(tmp, (size_t)len);
        }
      });
    }
    for (auto &t : th) t.join();
    for (auto &s : bufs)
      if (!s.empty()) XF_REQUIRE(fwrite(s.data(), 1, s.size(), fp) == s.size(), XF_EINVAL, "vtk: write failed");
  }
}

// device and pinned host staging of one chunk of node values
struct Staging {
  double *d = nullptr, *h = nullptr;
  ~Staging() {
    if (d) hipFree(d);
    if (h) hipHostFree(h);
  }
};

struct File {
  FILE *fp = nullptr;
  explicit File(const char *path, const char *mode) { fp = fopen(path, mode); }
  ~File() {
    if (fp) fclose(fp);
  }
};

}  // namespace

extern "C" int xf_vtk_write_flow(xf_ctx c, const char *path, const xf_vtk_flow *f) {
  return guarded(c, [&] {
    XF_REQUIRE(path && f && f->mask_grid_u && f->mask_grid_p && f->mask_grid_d && f->mask_grid_c && f->x && f->y &&
                   f->z && f->u && f->v && f->w,
               XF_EINVAL, "xf_vtk_write_flow: null argument");
    const xf_grid gs[4] = {f->mask_grid_u, f->mask_grid_p, f->mask_grid_d, f->mask_grid_c};
    const int n = gs[0]->n, l = gs[0]->l, m = gs[0]->m;
    for (auto g : gs)
      XF_REQUIRE(g->n == n && g->l == l && g->m == m, XF_EINVAL, "xf_vtk_write_flow: the grids differ in size");
    XF_REQUIRE(n >= 3 && l >= 3 && m >= 3, XF_EINVAL, "xf_vtk_write_flow: fewer than 3 nodes along an axis");
    const int64_t nodes = (int64_t)(n - 2) * (l - 2) * (m - 2);
    XF_REQUIRE(nodes < ((int64_t)1 << 31), XF_EINVAL, "xf_vtk_write_flow: more than 2^31 - 1 nodes");
    File file(path, "w");
    XF_REQUIRE(file.fp, XF_EINVAL, "xf_vtk_write_flow: cannot open the output file");
    FILE *fp = file.fp;
    // XProblem.cpp:6142-6165
    fprintf(fp, "# vtk DataFile Version 3.0\n");
    fprintf(fp, "Flow\nASCII\n\n");
    fprintf(fp, "DATASET STRUCTURED_GRID\n");
    fprintf(fp, "DIMENSIONS %d %d %d\n", n - 2, l - 2, m - 2);
    fprintf(fp, "POINTS %d double\n", (int)nodes);
    const double *X = f->x, *Y = f->y, *Z = f->z;
    const int NI = n - 2, NJ = l - 2;
    format_parallel(nodes, [&](int64_t q, char *buf, size_t cap) {
      const int i = 1 + (int)(q % NI);
      const int64_t r = q / NI;
      const int j = 1 + (int)(r % NJ), k = 1 + (int)(r / NJ);
      return snprintf(buf, cap, "%f  %f  %f\n", X[i], Y[j], Z[k]);
    }, fp);
    Staging sg;
    XF_HIP(hipMalloc(&sg.d, sizeof(double) * 3 * kVtkChunk));
    XF_HIP(hipHostMalloc((void **)&sg.h, sizeof(double) * 3 * kVtkChunk, hipHostMallocDefault));
    double *dbuf = sg.d, *hbuf = sg.h;
    struct Pass {
      const char *head;
      const char *fmt;
      int what;
      xf_grid g;
      const double *a;
    };
    // XProblem.cpp:6170, 6202, 6245, 6290, 6335
    const Pass passes[5] = {
        {"POINT_DATA %d\n\nVECTORS Velocity double\n", nullptr, 0, gs[0], nullptr},
        {"\nSCALARS Pressure float\nLOOKUP_TABLE default\n", "%f\n ", 1, gs[1], f->p},
        {"\nSCALARS Density float\nLOOKUP_TABLE default\n", "%f\n", 2, gs[2], f->density},
        {"\nSCALARS Concentration float\nLOOKUP_TABLE default\n", "%f\n", 3, gs[3], f->concentration},
        {"\nSCALARS Viscosity float\nLOOKUP_TABLE default\n", "%f\n", 4, gs[3], f->viscosity},
    };
    for (const Pass &ps : passes) {
      if (ps.what == 0) fprintf(fp, ps.head, (int)nodes);
      else fprintf(fp, "%s", ps.head);
      for (int64_t q0 = 0; q0 < nodes; q0 += kVtkChunk) {
        const int64_t cnt = std::min(kVtkChunk, nodes - q0);
        const int nb = (int)std::min<int64_t>(4096, (cnt + 255) / 256);
        hipLaunchKernelGGL(k_vtk_nodes, dim3(nb), dim3(256), 0, c->stream, n, l, m, q0, cnt, ps.what, ps.g->d_codes,
                           ps.what == 0 ? f->u : ps.a, f->v, f->w, dbuf);
        check_kernel("k_vtk_nodes");
        const int per = ps.what == 0 ? 3 : 1;
        XF_HIP(hipMemcpyAsync(hbuf, dbuf, sizeof(double) * per * cnt, hipMemcpyDeviceToHost, c->stream));
        XF_HIP(hipStreamSynchronize(c->stream));
        if (ps.what == 0)
          format_parallel(cnt, [&](int64_t q, char *buf, size_t cap) {
            // "%f ", "%f ", "%f\n " (XProblem.cpp:6193-6195)
            return snprintf(buf, cap, "%f %f %f\n ", hbuf[3 * q], hbuf[3 * q + 1], hbuf[3 * q + 2]);
          }, fp);
        else
          format_parallel(cnt, [&](int64_t q, char *buf, size_t cap) { return snprintf(buf, cap, ps.fmt, hbuf[q]); },
                          fp);
      }
    }
    XF_REQUIRE(!ferror(fp), XF_EINVAL, "xf_vtk_write_flow: write failed");
  });
}

extern "C" int xf_vtk_write_boundary(xf_ctx c, const char *path, int circles, int in_circle, const double *xyz,
                                     const double *xyz0, const double *vel, const double *p, const double *wss,
                                     double time) {
  return guarded(c, [&] {
    XF_REQUIRE(path && xyz && xyz0 && vel && p && wss, XF_EINVAL, "xf_vtk_write_boundary: null argument");
    XF_REQUIRE(circles >= 1 && in_circle >= 1, XF_EINVAL, "xf_vtk_write_boundary: empty boundary");
    const int nn = circles * in_circle;
    // node arrays may live on the device or the host
    std::vector<double> hx(3 * (size_t)nn), hx0(3 * (size_t)nn), hv(3 * (size_t)nn), hp(nn), hs(nn);
    const double *src[5] = {xyz, xyz0, vel, p, wss};
    double *dst[5] = {hx.data(), hx0.data(), hv.data(), hp.data(), hs.data()};
    const size_t bytes[5] = {24 * (size_t)nn, 24 * (size_t)nn, 24 * (size_t)nn, 8 * (size_t)nn, 8 * (size_t)nn};
    for (int q = 0; q < 5; q++) XF_HIP(hipMemcpyAsync(dst[q], src[q], bytes[q], hipMemcpyDefault, c->stream));
    XF_HIP(hipStreamSynchronize(c->stream));
    (void)time;
    File file(path, "w");
    XF_REQUIRE(file.fp, XF_EINVAL, "xf_vtk_write_boundary: cannot open the output file");
    FILE *f = file.fp;
    // XProblem.cpp:9568-9632
    fprintf(f, "# vtk DataFile Version 1.0\nBound\nASCII\n\n");
    fprintf(f, "DATASET UNSTRUCTURED_GRID\n");
    fprintf(f, "POINTS %d double\n", nn);
    for (int q = 0; q < nn; q++) fprintf(f, "%f %f %f\n", hx[3 * q], hx[3 * q + 1], hx[3 * q + 2]);
    fprintf(f, "CELLS %d %d\n", in_circle * (circles - 1), 5 * in_circle * (circles - 1));
    for (int cr = 0; cr < circles - 1; ++cr)
      for (int i = 0; i < in_circle; ++i) {
        int i1 = cr * in_circle + i, i2 = i1 + 1, i3 = i2 + in_circle, i4 = i1 + in_circle;
        if (i == in_circle - 1) {  // the quad that closes the circle
          i2 = cr * in_circle;
          i3 = (cr + 1) * in_circle;
        }
        fprintf(f, "%d %d %d %d %d\n", 4, i1, i2, i3, i4);
      }
    fprintf(f, "CELL_TYPES %d\n", in_circle * (circles - 1));
    for (int q = 0; q < in_circle * (circles - 1); q++) fprintf(f, "%d\n", 9);
    fprintf(f, "POINT_DATA %d\n", nn);
    fprintf(f, "VECTORS Velocity double\n");
    for (int q = 0; q < nn; q++) fprintf(f, "%f %f %f\n", hv[3 * q], hv[3 * q + 1], hv[3 * q + 2]);
    fprintf(f, "VECTORS Displacement double\n");
    for (int q = 0; q < nn; q++) {
      const double dx = hx0[3 * q] - hx[3 * q], dy = hx0[3 * q + 1] - hx[3 * q + 1], dz = hx0[3 * q + 2] - hx[3 * q + 2];
      fprintf(f, "%f %f %f\n", dx, dy, dz);
    }
    fprintf(f, "\n\nSCALARS Pressure float\nLOOKUP_TABLE default\n");
    for (int q = 0; q < nn; q++) fprintf(f, "%f\n", hp[q]);
    fprintf(f, "\n\nSCALARS WSS float\nLOOKUP_TABLE default\n");
    for (int q = 0; q < nn; q++) fprintf(f, "%f\n", hs[q]);
    XF_REQUIRE(!ferror(f), XF_EINVAL, "xf_vtk_write_boundary: write failed");
  });
}

extern "C" int xf_npy_save(xf_ctx c, const char *path, const void *data, char kind, int word_size, int ndims,
                           const uint32_t *shape) {
  return guarded(c, [&] {
    XF_REQUIRE(path && data && shape && ndims >= 1, XF_EINVAL, "xf_npy_save: null argument");
    XF_REQUIRE((kind == 'f' && (word_size == 4 || word_size == 8)) ||
                   ((kind == 'i' || kind == 'u') && (word_size == 1 || word_size == 2 || word_size == 4 ||
                                                     word_size == 8)),
               XF_EINVAL, "xf_npy_save: unsupported element type");
    // header: cnpy.h:207-236 (little-endian descr, shape "(a, b, c)" with a
    // trailing comma for one dimension, dictionary padded with spaces so that
    // 10 + its length is a multiple of 16, ending in a newline)
    std::string dict = "{'descr': '<";
    dict += kind;
    dict += std::to_string(word_size);
    dict += "', 'fortran_order': False, 'shape': (";
    size_t count = 1;
    for (int d = 0; d < ndims; d++) {
      if (d) dict += ", ";
      dict += std::to_string(shape[d]);
      count *= shape[d];
    }
    if (ndims == 1) dict += ",";
    dict += "), }";
    const int remainder = 16 - (int)((10 + dict.size()) % 16);
    dict.append((size_t)remainder, ' ');
    dict.back() = '\n';
    XF_REQUIRE(dict.size() < 65536, XF_EINVAL, "xf_npy_save: header too long");
    std::string header;
    header += (char)0x93;
    header += "NUMPY";
    header += (char)0x01;
    header += (char)0x00;
    const uint16_t hl = (uint16_t)dict.size();
    header += (char)(hl & 0xff);
    header += (char)(hl >> 8);
    header += dict;
    File file(path, "wb");
    XF_REQUIRE(file.fp, XF_EINVAL, "xf_npy_save: cannot open the output file");
    XF_REQUIRE(fwrite(header.data(), 1, header.size(), file.fp) == header.size(), XF_EINVAL, "xf_npy_save: write failed");
    // data: from host or device memory, in pieces of 64 MB
    const size_t total = count * (size_t)word_size, piece = (size_t)64 << 20;
    std::vector<char> h(std::min(total, piece));
    for (size_t off = 0; off < total; off += piece) {
      const size_t len = std::min(piece, total - off);
      XF_HIP(hipMemcpyAsync(h.data(), (const char *)data + off, len, hipMemcpyDefault, c->stream));
      XF_HIP(hipStreamSynchronize(c->stream));
      XF_REQUIRE(fwrite(h.data(), 1, len, file.fp) == len, XF_EINVAL, "xf_npy_save: write failed");
    }
  });
}
