// Learned preconditioner: a 3-level 3-D U-Net (channels 16/32/64, 3x3x3
// convolutions, BF16 weights and activations, FP32 accumulation), with
// hand-written forward, backward and Adam kernels.
//
//   L1 (S):    a0 = relu(c0(r, mask))  2->16      a1 = relu(c1(a0))  16->16
//   L2 (S/2):  a2 = relu(c2(pool a1))  16->32     a3 = relu(c3(a2))  32->32
//   L3 (S/4):  a4 = relu(c4(pool a3))  32->64     a5 = relu(c5(a4))  64->64
//   L2:        a6 = relu(c6([up a5, a3])) 96->32  a7 = relu(c7(a6))  32->32
//   L1:        a8 = relu(c8([up a7, a1])) 48->16  y = c9(a8)         16->1
// pool = 2x2x2 average, up = 2x nearest, [.,.] = channel concatenation.
//
// Tensors are channels-last ([b][z][y][x][c]).  c1..c8 run on one implicit-GEMM
// kernel: D[co][cell] = sum_k W[co][k] X[k][cell], k = tap*Cin + ci, with
// v_mfma_f32_16x16x32_bf16 (16 output channels x 16 cells x 32 k per
// instruction); the block stages its input tile (+1 halo) in LDS, each wave
// keeps the weight fragments of a k-step in registers and applies them to
// several rows of 16 cells.  The input-gradient of a convolution is the same
// kernel with the weights flipped and transposed.  c0, c9, the weight
// gradients and the element-wise steps are plain FP32 kernels.
//
// As a preconditioner the network maps r to e ~ A^-1 r on the interior of a
// pressure grid: the input is scaled by s = rms(r) over the unknowns and the
// output multiplied by s (so M(a r) = a M(r)), and masked to the unknowns.
#include <hip/hip_bf16.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "xf_internal.h"

typedef __bf16 bf16;
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

constexpr int NL = 10;  // layers c0..c9
static const int LCIN[NL] = {2, 16, 16, 32, 32, 64, 96, 32, 48, 16};
static const int LCOUT[NL] = {16, 16, 32, 32, 64, 64, 32, 32, 16, 1};

// ---------------------------------------------------------------- helpers
__device__ inline bf16 to_bf16(float f) { return (bf16)f; }
__device__ inline float from_bf16(bf16 h) { return (float)h; }

// K of a convolution: 27 taps of Cin channels, zero-padded to a multiple of 96
// (k-steps of 32, unrolled by 3)
static int kpad_of(int cin) { return (27 * cin + 95) / 96 * 96; }

// ------------------------------------------------------ MFMA convolution
struct ConvArgs {
  const bf16 *src0;  // [B][D>>up0][H>>up0][W>>up0][C0]
  const bf16 *src1;  // [B][D][H][W][C1] or null
  const bf16 *wt;    // [Cout][Kpad]
  const float *bias; // [Cout] or null
  bf16 *out;         // [B][D][H][W][Cout]
  int C0, C1, up0, Cout, D, H, W, relu, TZ, TY, Kpad;
  int nsplit;        // blocks per spatial tile, each taking NT of the Cout/16 channel tiles
  uint32_t m_cpc, m_sy;  // division magics for Cin / 8 and TY + 2
};

// q / d for q * d <= 2^32 as a multiply-high by m = ceil(2^32 / d)
__device__ inline int fdiv(int q, uint32_t m) { return (int)__umulhi((uint32_t)q, m); }
static uint32_t fdiv_magic(int d) { return (uint32_t)((((uint64_t)1 << 32) + (uint64_t)d - 1) / (uint64_t)d); }

// One block: a 16 x TY x TZ cell tile (RPW rows of 16 cells per wave, four
// waves) and NT 16-channel output tiles.  Small levels (a 16^3 volume is 256
// rows of 16 cells) split Cout over nsplit blocks and use short tiles so that
// the launch still covers the chip; every output sums the same k-steps in the
// same order whatever the split.  The k-loop walks (tap, channel) with
// uniform (scalar) counters; only the lane's cell and channel group are
// per-lane.
template <int NT, int RPW>
__global__ __launch_bounds__(256) void k_conv_mfma(ConvArgs a) {
  extern __shared__ __attribute__((aligned(16))) bf16 tile[];
  const int Cin = a.C0 + a.C1;
  const int CS = Cin + 8;  // LDS cell stride (bf16), 16-byte multiple
  constexpr int TX = 16, SX = TX + 2;
  const int SY = a.TY + 2, SZ = a.TZ + 2;
  const int xt = a.W / TX, yt = a.H / a.TY, zt = a.D / a.TZ;
  int bid = blockIdx.x;
  const int ct0 = (bid % a.nsplit) * NT;  // first output channel tile
  bid /= a.nsplit;
  const int bx = bid % xt; bid /= xt;
  const int by = bid % yt; bid /= yt;
  const int bz = bid % zt; bid /= zt;
  const int b = bid;
  const int x0 = bx * TX, y0 = by * a.TY, z0 = bz * a.TZ;
  // stage the input tile with a one-cell halo (zero outside the volume)
  const int cpc = Cin / 8;  // 16-byte chunks per cell
  const int nchunk = SZ * SY * SX * cpc;
  const int D0 = a.D >> a.up0, H0 = a.H >> a.up0, W0 = a.W >> a.up0;
  // up to twelve chunks per thread in flight: loads first, then the LDS stores
  // (a 48-channel 2 x 4 tile is 10 chunks per thread: one round trip)
  constexpr int U = 12;
  for (int qb = threadIdx.x; qb < nchunk; qb += U * blockDim.x) {
    bf16x8 v[U];
    int dst[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int q = qb + u * blockDim.x;
      v[u] = bf16x8{};
      dst[u] = -1;
      if (q < nchunk) {
        const int cell = fdiv(q, a.m_cpc), cc = (q - cell * cpc) * 8;
        const int row = cell / SX, xx = cell - row * SX;
        const int zz = fdiv(row, a.m_sy), yy = row - zz * SY;
        const int x = x0 - 1 + xx, y = y0 - 1 + yy, z = z0 - 1 + zz;
        dst[u] = cell * CS + cc;
        if (x >= 0 && x < a.W && y >= 0 && y < a.H && z >= 0 && z < a.D) {
          if (cc < a.C0) {
            const int64_t g =
                ((((int64_t)b * D0 + (z >> a.up0)) * H0 + (y >> a.up0)) * W0 + (x >> a.up0)) * a.C0 + cc;
            v[u] = *(const bf16x8 *)(a.src0 + g);
          } else {
            const int64_t g = ((((int64_t)b * a.D + z) * a.H + y) * a.W + x) * a.C1 + (cc - a.C0);
            v[u] = *(const bf16x8 *)(a.src1 + g);
          }
        }
      }
    }
#pragma unroll
    for (int u = 0; u < U; u++)
      if (dst[u] >= 0) *(bf16x8 *)(tile + dst[u]) = v[u];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int l16 = lane & 15, lg = lane >> 4;
  // K = 27 taps x Cin channels, k = tap * Cin + ci, zero-padded to Kpad (a
  // multiple of 96, so the unroll by 3 below divides the step count).  Lane
  // group lg takes k = 32 s + 8 lg + j of step s: 8 channels of one tap.  A
  // step spans at most two taps (Cin >= 16); the walk over them is uniform
  // (scalar): the step starts at channel c0 of tap `tap`, whose LDS offset is
  // off_lo, and off_hi is the next tap's.  Taps past the last read tap 26
  // (finite values against zero weights).
  int vcell[RPW];
#pragma unroll
  for (int r = 0; r < RPW; r++) {
    const int rho = wave + 4 * r, rz = rho / a.TY, ry = rho - rz * a.TY;
    vcell[r] = ((rz * SY + ry) * SX + l16) * CS;
  }
  f32x4 acc[RPW][NT];
#pragma unroll
  for (int r = 0; r < RPW; r++)
#pragma unroll
    for (int t = 0; t < NT; t++) acc[r][t] = f32x4{0, 0, 0, 0};
  // off_hi follows tap + 1 by its digits (hx, hy) while tap + 1 <= 26 and
  // then stays at tap 26; the offset step between consecutive taps is CS,
  // plus jx where x wraps, plus jy where y wraps too.
  const int jx = (SX - 3) * CS, jy = (SY - 3) * SX * CS;
  int tap = 0, c0 = 0, off_lo = 0, off_hi = CS, hx = 1, hy = 0;
  const int k8 = 8 * lg;
  auto advance = [&]() {
    c0 += 32;
    while (c0 >= Cin) {
      c0 -= Cin;
      ++tap;
      off_lo = off_hi;
      if (tap + 1 <= 26) {
        int inc = CS;
        if (++hx == 3) {
          hx = 0;
          inc += jx;
          if (++hy == 3) {
            hy = 0;
            inc += jy;
          }
        }
        off_hi += inc;
      }
    }
  };
  auto load_b = [&](bf16x8 *bf) {
    const int k0 = c0 + k8;
    const int o = k0 < Cin ? off_lo + k0 : off_hi + (k0 - Cin);
#pragma unroll
    for (int r = 0; r < RPW; r++) bf[r] = *(const bf16x8 *)(tile + vcell[r] + o);
  };
  // Weight fragments come from global memory (L2-resident: every block reads
  // the same weights), packed in fragment order (one k-step of a 16-channel
  // output tile is 1 KB, lane L's 16 bytes at 16 L), and fetched WP k-steps
  // ahead into a ring of registers indexed at compile time (the k-loop is
  // unrolled by WP, which divides Kpad / 32), so an L2 round trip is paid once
  // per WP steps; input fragments (LDS) are loaded one step ahead.  Fetches
  // past the last step repeat the last one.
  constexpr int WP = 3;
  const int nsteps = a.Kpad / 32;
  const bf16 *wbase = a.wt + (int64_t)ct0 * 16 * a.Kpad + lane * 8;
  const int wstride = 16 * a.Kpad;
  // Both rings are indexed at compile time by the unrolled step u, so no
  // fragment is ever copied between registers: step u multiplies the weights
  // aw[u] and the inputs bq[u], then refills aw[u] with step u + WP's weights
  // (the MFMA has read them at issue) and bq[(u + 1) % WP] with step u + 1's
  // inputs.  (Copying fragments through a "current" register set cost about
  // twenty vector instructions per MFMA: the loop was issue-bound.)
  bf16x8 aw[WP][NT];
#pragma unroll
  for (int u = 0; u < WP; u++)
#pragma unroll
    for (int t = 0; t < NT; t++) aw[u][t] = *(const bf16x8 *)(wbase + u * 512 + t * wstride);
  bf16x8 bq[WP][RPW];
  load_b(bq[0]);
  for (int s0 = 0; s0 < nsteps; s0 += WP) {
#pragma unroll
    for (int u = 0; u < WP; u++) {
      const int s = s0 + u;
      advance();
      load_b(bq[(u + 1) % WP]);
#pragma unroll
      for (int r = 0; r < RPW; r++)
#pragma unroll
        for (int t = 0; t < NT; t++)
          acc[r][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aw[u][t], bq[u][r], acc[r][t], 0, 0, 0);
      const bf16 *wq = wbase + min(s + WP, nsteps - 1) * 512;
#pragma unroll
      for (int t = 0; t < NT; t++) aw[u][t] = *(const bf16x8 *)(wq + t * wstride);
    }
  }
  // epilogue: lane holds D[co = 16t + 4 lg + e][cell = l16]
#pragma unroll
  for (int r = 0; r < RPW; r++) {
    const int rho = wave + 4 * r, rz = rho / a.TY, ry = rho - rz * a.TY;
    const int64_t cell = (((int64_t)b * a.D + z0 + rz) * a.H + y0 + ry) * a.W + x0 + l16;
#pragma unroll
    for (int t = 0; t < NT; t++) {
      const int co = (ct0 + t) * 16 + 4 * lg;
      bf16 o[4];
      for (int e = 0; e < 4; e++) {
        float v = acc[r][t][e] + (a.bias ? a.bias[co + e] : 0.0f);
        if (a.relu) v = fmaxf(v, 0.0f);
        o[e] = to_bf16(v);
      }
      // 4 consecutive channels: one 8-byte store
      uint64_t packed;
      memcpy(&packed, o, 8);
      *(uint64_t *)(a.out + cell * a.Cout + co) = packed;
    }
  }
}

// ------------------------------------------------- first and last layers
// Input source of c0: value(b, z, y, x) = r[...] * inv_s and mask[...], read
// either from crop tensors [B][S][S][S] (ring = 0) or from a solver field with
// a one-cell ring (ring = 1, mask = unknown flag of the grid).
struct Src0 {
  const float *r;
  const float *mask;        // crop mask or null
  const uint8_t *flags;     // solver flags or null
  const double *scale;      // device scalar: sum r^2 (solver mode) or null
  double count;             // unknowns (solver mode)
  int ring, D, H, W;
};

__device__ inline void src0_read(const Src0 &s, int b, int z, int y, int x, float inv, float &rv, float &mv) {
  if (x < 0 || y < 0 || z < 0 || x >= s.W || y >= s.H || z >= s.D) {
    rv = 0;
    mv = 0;
    return;
  }
  if (s.ring) {
    const int64_t c = (((int64_t)(z + 1) * (s.H + 2)) + (y + 1)) * (s.W + 2) + (x + 1);
    mv = (s.flags[c] & FL_UNKNOWN) ? 1.f : 0.f;
    rv = s.r[c] * inv * mv;
  } else {
    const int64_t c = (((int64_t)b * s.D + z) * s.H + y) * s.W + x;
    mv = s.mask[c];
    rv = s.r[c] * inv;
  }
}

__device__ inline float src0_inv(const Src0 &s) {
  if (!s.scale) return 1.0f;
  const double ms = *s.scale / (s.count > 0 ? s.count : 1.0);
  return ms > 0 ? (float)(1.0 / sqrt(ms)) : 0.0f;
}

// c0 (2 -> 16, FP32 with fused multiply-adds): a block computes a 16 x 4 x 4
// cell tile from its input (r * inv_s, mask) staged in LDS with a one-cell halo.
__global__ __launch_bounds__(256) void k_conv_first(Src0 s, const float *__restrict__ w,
                                                    const float *__restrict__ bias, bf16 *out, int B) {
  // w: [16][27][2] fp32 master weights
  constexpr int SX = 18, SY = 6, SZ = 6;
  __shared__ float2 in[SZ * SY * SX];
  const float inv = src0_inv(s);
  const int xt = s.W / 16, yt = s.H / 4, zt = s.D / 4;
  int bid = blockIdx.x;
  const int bx = bid % xt; bid /= xt;
  const int by = bid % yt; bid /= yt;
  const int bz = bid % zt;
  const int b = bid / zt;
  const int x0 = bx * 16, y0 = by * 4, z0 = bz * 4;
  // all of a thread's cells (at most 3) read before the LDS stores: one round trip
  constexpr int NQ = SZ * SY * SX, U = (NQ + 255) / 256;
  float2 v[U];
#pragma unroll
  for (int u = 0; u < U; u++) {
    const int q = threadIdx.x + u * 256;
    v[u] = make_float2(0.f, 0.f);
    if (q < NQ) {
      const int row = q / SX, xx = q - row * SX, zz = row / SY, yy = row - zz * SY;
      src0_read(s, b, z0 - 1 + zz, y0 - 1 + yy, x0 - 1 + xx, inv, v[u].x, v[u].y);
    }
  }
#pragma unroll
  for (int u = 0; u < U; u++) {
    const int q = threadIdx.x + u * 256;
    if (q < NQ) in[q] = v[u];
  }
  __syncthreads();
  const int tx = threadIdx.x & 15, ty = (threadIdx.x >> 4) & 3, tz = threadIdx.x >> 6;
  float acc[16];
#pragma unroll
  for (int co = 0; co < 16; co++) acc[co] = bias[co];
#pragma unroll
  for (int t = 0; t < 27; t++) {
    const float2 v = in[((tz + t / 9) * SY + (ty + (t / 3) % 3)) * SX + tx + t % 3];
#pragma unroll
    for (int co = 0; co < 16; co++) acc[co] = fmaf(w[(co * 27 + t) * 2 + 1], v.y, fmaf(w[(co * 27 + t) * 2], v.x, acc[co]));
  }
  bf16 o[16];
#pragma unroll
  for (int co = 0; co < 16; co++) o[co] = to_bf16(fmaxf(acc[co], 0.0f));
  const int64_t q = (((int64_t)b * s.D + z0 + tz) * s.H + y0 + ty) * s.W + x0 + tx;
  for (int co = 0; co < 16; co += 8) *(bf16x8 *)(out + q * 16 + co) = *(bf16x8 *)(o + co);
}

// Output of c9: y = bias + sum w * a8 (fp32).  Crop mode writes y[B][S^3];
// solver mode writes z[ring layout] = s * y * mask and zero on the ring.
struct Dst9 {
  float *y;
  const uint8_t *flags;
  const double *scale;
  double count;
  int ring;
};

__global__ void k_conv_last(const bf16 *__restrict__ a8, const float *__restrict__ w, const float *__restrict__ bias,
                            Dst9 d, int B, int D, int H, int W) {
  const int64_t n = (int64_t)B * D * H * W;
  float s = 1.0f;
  if (d.scale) {
    const double ms = *d.scale / (d.count > 0 ? d.count : 1.0);
    s = (float)sqrt(ms > 0 ? ms : 0.0);
  }
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    const unsigned qq = (unsigned)q, qw = qq / (unsigned)W, qh = qw / (unsigned)H;
    const int x = (int)(qq - qw * W), y = (int)(qw - qh * H), z = (int)(qh % (unsigned)D), b = (int)(qh / D);
    float acc = bias[0];
    for (int t = 0; t < 27; t++) {
      const int zz = z + t / 9 - 1, yy = y + (t / 3) % 3 - 1, xx = x + t % 3 - 1;
      if (xx < 0 || yy < 0 || zz < 0 || xx >= W || yy >= H || zz >= D) continue;
      const bf16 *p = a8 + ((((int64_t)b * D + zz) * H + yy) * W + xx) * 16;
      for (int ci = 0; ci < 16; ci++) acc += w[t * 16 + ci] * from_bf16(p[ci]);
    }
    if (d.ring) {
      const int64_t c = (((int64_t)(z + 1) * (H + 2)) + (y + 1)) * (W + 2) + (x + 1);
      d.y[c] = (d.flags[c] & FL_UNKNOWN) ? acc * s : 0.0f;
    } else {
      d.y[q] = acc;
    }
  }
}

// c9 on a block of 16 x 4 x 4 cells with their one-cell halo of a8 staged in
// LDS (k_conv_last reads the 27 neighbour cells, 32 bytes each, straight from
// memory for every cell); the same sums in the same order.  W % 16, H % 4 and
// D % 4 must be 0.
__global__ __launch_bounds__(256) void k_conv_last_tile(const bf16 *__restrict__ a8, const float *__restrict__ w,
                                                        const float *__restrict__ bias, Dst9 d, int B, int D, int H,
                                                        int W) {
  constexpr int SX = 18, SY = 6, SZ = 6, CS = 24;  // LDS cell stride: 16 channels + 8 pad (48 bytes)
  __shared__ __attribute__((aligned(16))) bf16 tile[SZ * SY * SX * CS];
  const int xt = W / 16, yt = H / 4, zt = D / 4;
  int bid = blockIdx.x;
  const int bx = bid % xt; bid /= xt;
  const int by = bid % yt; bid /= yt;
  const int bz = bid % zt;
  const int b = bid / zt;
  const int x0 = bx * 16, y0 = by * 4, z0 = bz * 4;
  // all of a thread's chunks (at most 6) loaded before the LDS stores: one round trip
  constexpr int NQ = SZ * SY * SX * 2, U = (NQ + 255) / 256;
  bf16x8 v[U];
#pragma unroll
  for (int u = 0; u < U; u++) {
    const int q = threadIdx.x + u * 256;
    v[u] = bf16x8{};
    if (q < NQ) {
      const int cell = q >> 1, half = q & 1;
      const int row = cell / SX, xx = cell - row * SX, zz = row / SY, yy = row - zz * SY;
      const int x = x0 - 1 + xx, y = y0 - 1 + yy, z = z0 - 1 + zz;
      if (x >= 0 && y >= 0 && z >= 0 && x < W && y < H && z < D)
        v[u] = *(const bf16x8 *)(a8 + ((((int64_t)b * D + z) * H + y) * W + x) * 16 + half * 8);
    }
  }
#pragma unroll
  for (int u = 0; u < U; u++) {
    const int q = threadIdx.x + u * 256;
    if (q < NQ) *(bf16x8 *)(tile + (q >> 1) * CS + (q & 1) * 8) = v[u];
  }
  __syncthreads();
  const int tx = threadIdx.x & 15, ty = (threadIdx.x >> 4) & 3, tz = threadIdx.x >> 6;
  const int x = x0 + tx, y = y0 + ty, z = z0 + tz;
  float s = 1.0f;
  if (d.scale) {
    const double ms = *d.scale / (d.count > 0 ? d.count : 1.0);
    s = (float)sqrt(ms > 0 ? ms : 0.0);
  }
  float acc = bias[0];
#pragma unroll
  for (int t = 0; t < 27; t++) {
    const int dz = t / 9, dy = (t / 3) % 3, dx = t % 3;
    const int zz = z + dz - 1, yy = y + dy - 1, xx = x + dx - 1;
    if (xx < 0 || yy < 0 || zz < 0 || xx >= W || yy >= H || zz >= D) continue;
    const bf16 *p = tile + (((tz + dz) * SY + (ty + dy)) * SX + tx + dx) * CS;
    bf16 v[16];
    *(bf16x8 *)v = *(const bf16x8 *)p;
    *(bf16x8 *)(v + 8) = *(const bf16x8 *)(p + 8);
#pragma unroll
    for (int ci = 0; ci < 16; ci++) acc += w[t * 16 + ci] * from_bf16(v[ci]);
  }
  if (d.ring) {
    const int64_t c = (((int64_t)(z + 1) * (H + 2)) + (y + 1)) * (W + 2) + (x + 1);
    d.y[c] = (d.flags[c] & FL_UNKNOWN) ? acc * s : 0.0f;
  } else {
    d.y[(((int64_t)b * D + z) * H + y) * W + x] = acc;
  }
}

// ------------------------------------------------------------ elementwise
__global__ void k_pool(const bf16 *__restrict__ in, bf16 *__restrict__ out, int B, int D, int H, int W, int C) {
  // in [B][D][H][W][C] -> out [B][D/2][H/2][W/2][C], 2x2x2 mean
  const int d = D / 2, h = H / 2, w = W / 2;
  const int64_t n = (int64_t)B * d * h * w * C;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(q % C);
    int64_t cell = q / C;
    const int x = (int)(cell % w), y = (int)((cell / w) % h), z = (int)((cell / ((int64_t)w * h)) % d);
    const int b = (int)(cell / ((int64_t)w * h * d));
    float s = 0;
    for (int t = 0; t < 8; t++) {
      const int zz = 2 * z + (t >> 2), yy = 2 * y + ((t >> 1) & 1), xx = 2 * x + (t & 1);
      s += from_bf16(in[((((int64_t)b * D + zz) * H + yy) * W + xx) * C + c]);
    }
    out[q] = to_bf16(s * 0.125f);
  }
}

// g (bf16, C channels) *= (act > 0)
__global__ void k_relu_back(bf16 *__restrict__ g, const bf16 *__restrict__ act, int64_t n) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
    if (!(from_bf16(act[q]) > 0.0f)) g[q] = to_bf16(0.0f);
}

// coarse[b][z][y][x][c] (+)= sum over the 8 children of fine[..][c0 + c]
// (backward of nearest upsampling taking channels [c0, c0+C) of a Cf-channel tensor)
__global__ void k_upsample_back(const bf16 *__restrict__ fine, int Cf, int c0, bf16 *__restrict__ coarse, int C, int B,
                                int D, int H, int W) {
  const int64_t n = (int64_t)B * D * H * W * C;  // coarse sizes
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(q % C);
    int64_t cell = q / C;
    const int x = (int)(cell % W), y = (int)((cell / W) % H), z = (int)((cell / ((int64_t)W * H)) % D);
    const int b = (int)(cell / ((int64_t)W * H * D));
    float s = 0;
    for (int t = 0; t < 8; t++) {
      const int zz = 2 * z + (t >> 2), yy = 2 * y + ((t >> 1) & 1), xx = 2 * x + (t & 1);
      s += from_bf16(fine[((((int64_t)b * 2 * D + zz) * 2 * H + yy) * 2 * W + xx) * Cf + c0 + c]);
    }
    coarse[q] = to_bf16(s);
  }
}

// fine grad [..][C] = slice[..][c0 + c] + (coarse pooled grad)/8 at the parent
// (the skip tensor's gradient: concat slice plus average-pool backward)
__global__ void k_skip_grad(const bf16 *__restrict__ slice, int Cs, int c0, const bf16 *__restrict__ pooled,
                            bf16 *__restrict__ out, int C, int B, int D, int H, int W) {
  const int64_t n = (int64_t)B * D * H * W * C;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(q % C);
    int64_t cell = q / C;
    const int x = (int)(cell % W), y = (int)((cell / W) % H), z = (int)((cell / ((int64_t)W * H)) % D);
    const int b = (int)(cell / ((int64_t)W * H * D));
    const int64_t pc = ((((int64_t)b * (D / 2) + z / 2) * (H / 2) + y / 2) * (W / 2) + x / 2) * C + c;
    out[q] = to_bf16(from_bf16(slice[cell * Cs + c0 + c]) + from_bf16(pooled[pc]) * 0.125f);
  }
}

// ------------------------------------------------------- weight gradients
// dW[co][tap][ci] += sum_cells g[cell][co] * X[cell + tap][ci], db[co] += sum g.
// X is the (optionally two-source, upsampled) input as in the forward pass.
// One block per (co, tap) x cell-chunk; fp32 partial sums, then atomicAdd.
struct GradArgs {
  const bf16 *src0;
  const bf16 *src1;
  int C0, C1, up0;
  const bf16 *g;     // [B][D][H][W][Cout] (bf16) or null
  const float *g32;  // fp32 single-channel gradient (c9) or null
  int Cout, D, H, W, B;
  float *dw, *db;
};

// 8 consecutive input channels (ci8 .. ci8+7) of the (two-source, upsampled) input
__device__ inline void x_read8(const GradArgs &a, int b, int z, int y, int x, int ci8, float v[8]) {
  if (x < 0 || y < 0 || z < 0 || x >= a.W || y >= a.H || z >= a.D) {
    for (int e = 0; e < 8; e++) v[e] = 0;
    return;
  }
  bf16x8 h;
  if (ci8 < a.C0) {
    const int D0 = a.D >> a.up0, H0 = a.H >> a.up0, W0 = a.W >> a.up0;
    h = *(const bf16x8 *)(a.src0 +
                          ((((int64_t)b * D0 + (z >> a.up0)) * H0 + (y >> a.up0)) * W0 + (x >> a.up0)) * a.C0 + ci8);
  } else {
    h = *(const bf16x8 *)(a.src1 + ((((int64_t)b * a.D + z) * a.H + y) * a.W + x) * a.C1 + (ci8 - a.C0));
  }
  for (int e = 0; e < 8; e++) v[e] = from_bf16(h[e]);
}

// grid (chunks, Cout, 27 * Cin/8): each block accumulates dW[co][tap][ci8..+8]
// over a chunk of cells; the (tap 13, first channel group) block also sums db.
__global__ __launch_bounds__(256) void k_conv_wgrad(GradArgs a, int chunks) {
  __shared__ float red[9][256];
  const int Cin = a.C0 + a.C1;
  const int co = blockIdx.y;
  const int tap = blockIdx.z / (Cin / 8), ci8 = (blockIdx.z % (Cin / 8)) * 8;
  const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
  const int64_t ncell = (int64_t)a.B * a.D * a.H * a.W;
  const int64_t per = (ncell + chunks - 1) / chunks;
  const int64_t c_beg = blockIdx.x * per, c_end = min(ncell, c_beg + per);
  float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  float gb = 0;
  for (int64_t q = c_beg + threadIdx.x; q < c_end; q += blockDim.x) {
    const float gv = a.g32 ? a.g32[q] : from_bf16(a.g[q * a.Cout + co]);
    if (gv == 0.0f) continue;
    const int x = (int)(q % a.W), y = (int)((q / a.W) % a.H), z = (int)((q / ((int64_t)a.W * a.H)) % a.D);
    const int b = (int)(q / ((int64_t)a.W * a.H * a.D));
    gb += gv;
    float v[8];
    x_read8(a, b, z + dz, y + dy, x + dx, ci8, v);
    for (int e = 0; e < 8; e++) acc[e] += gv * v[e];
  }
  for (int e = 0; e < 8; e++) red[e][threadIdx.x] = acc[e];
  red[8][threadIdx.x] = gb;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st)
      for (int e = 0; e < 9; e++) red[e][threadIdx.x] += red[e][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x < 8) atomicAdd(&a.dw[((int64_t)co * 27 + tap) * Cin + ci8 + threadIdx.x], red[threadIdx.x][0]);
  if (threadIdx.x == 8 && tap == 13 && ci8 == 0) atomicAdd(&a.db[co], red[8][0]);
}

// input gradient of c9 (16 channels) from the fp32 output gradient dy
__global__ void k_last_dx(const float *__restrict__ dy, const float *__restrict__ w, bf16 *__restrict__ g8, int B, int D,
                          int H, int W) {
  const int64_t n = (int64_t)B * D * H * W;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    const unsigned qq = (unsigned)q, qw = qq / (unsigned)W, qh = qw / (unsigned)H;
    const int x = (int)(qq - qw * W), y = (int)(qw - qh * H), z = (int)(qh % (unsigned)D), b = (int)(qh / D);
    float acc[16] = {};
    for (int t = 0; t < 27; t++) {
      // output cell o = q - offset(t) contributes w[t] * dy[o] to this input cell
      const int zz = z - (t / 9 - 1), yy = y - ((t / 3) % 3 - 1), xx = x - (t % 3 - 1);
      if (xx < 0 || yy < 0 || zz < 0 || xx >= W || yy >= H || zz >= D) continue;
      const float g = dy[(((int64_t)b * D + zz) * H + yy) * W + xx];
      for (int ci = 0; ci < 16; ci++) acc[ci] += g * w[t * 16 + ci];
    }
    for (int ci = 0; ci < 16; ci++) g8[q * 16 + ci] = to_bf16(acc[ci]);
  }
}

// Packed weights are in MFMA fragment order: for 16-row tile rt, k-step s
// (k = 32 s .. 32 s + 31) and lane L = 16 g + i, the 8 values at
// ((rt * Kpad / 32 + s) * 64 + L) * 8 are rows 16 rt + i, k = 32 s + 8 g + j,
// j = 0..7 (zero past the rows or K).  A wave's fragment of one step is then
// 1 KB contiguous.
__device__ inline void frag_index(int64_t q, int Kpad, int &row, int &k) {
  const int64_t tile = (int64_t)16 * Kpad;
  const int rt = (int)(q / tile);
  const int rem = (int)(q - rt * tile);
  const int s = rem >> 9, lane = (rem >> 3) & 63, j = rem & 7;
  row = rt * 16 + (lane & 15);
  k = 32 * s + 8 * (lane >> 4) + j;
}

// Input-gradient weights: Wt'[ci][k' = tap' * Cout + co] = W[co][26 - tap'][ci]
// (bf16, fragment order)
__global__ void k_flip_weights(const float *__restrict__ w, int Cout, int Cin, bf16 *__restrict__ wt, int Kpad) {
  const int64_t n = (int64_t)Cin * Kpad;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    int ci, k;
    frag_index(q, Kpad, ci, k);
    float v = 0;
    if (ci < Cin && k < 27 * Cout) {
      const int tp = k / Cout, co = k - tp * Cout;
      v = w[((int64_t)co * 27 + (26 - tp)) * Cin + ci];
    }
    wt[q] = to_bf16(v);
  }
}

// forward weights: bf16 W[co][k = tap * Cin + ci] (fragment order) from fp32 [Cout][27][Cin]
__global__ void k_pack_weights(const float *__restrict__ w, int Cout, int Cin, bf16 *__restrict__ wt, int Kpad) {
  const int64_t n = (int64_t)Cout * Kpad;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    int co, k;
    frag_index(q, Kpad, co, k);
    wt[q] = to_bf16(co < Cout && k < 27 * Cin ? w[(int64_t)co * 27 * Cin + k] : 0.0f);
  }
}

__global__ void k_sum_f32(const float *__restrict__ v, int64_t n, double *partials, unsigned *counter, double *out) {
  __shared__ double lds[16];
  double acc = 0;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
    acc += v[q];
  double bs = block_sum(acc, lds);
  double tot;
  if (grid_sum_last(bs, partials, counter, &tot, lds) && threadIdx.x == 0) *out = tot;
}

// masked mean-squared error, dy = 2 mask (y - t) / N_mask (N_mask from *count)
__global__ void k_mse(const float *__restrict__ y, const float *__restrict__ t, const float *__restrict__ mask, int64_t n,
                      float *__restrict__ dy, double *partials, unsigned *counter, double *out, const double *count) {
  __shared__ double lds[16];
  const double inv_count = *count > 0 ? 1.0 / *count : 0.0;
  double acc = 0;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    const float d = (y[q] - t[q]) * mask[q];
    acc += (double)d * d;
    dy[q] = (float)(2.0 * d * inv_count);
  }
  double bs = block_sum(acc, lds);
  double tot;
  if (grid_sum_last(bs, partials, counter, &tot, lds) && threadIdx.x == 0) *out = tot * inv_count;
}

// Scale-invariant loss per sample b: L_b = 1 - (y,t)^2 / ((y,y)(t,t)) over
// the masked cells, the loss that matters to a flexible-CG step (which
// rescales the preconditioned residual itself); the batch loss is the mean.
// Three deterministic passes: per-(sample, chunk) partial dots, per-sample
// coefficients, then dL/dy = -(2 (y,t) / ((y,y)(t,t) B)) (t - ((y,t)/(y,y)) y).
constexpr int kCosChunks = 64;

__global__ void k_cos_dots(const float *__restrict__ y, const float *__restrict__ t, const float *__restrict__ mask,
                           int64_t per_sample, double *__restrict__ part) {
  __shared__ double lds[16];
  const int b = blockIdx.y, ch = blockIdx.x;
  const int64_t chunk = (per_sample + kCosChunks - 1) / kCosChunks;
  const int64_t beg = ch * chunk, end = min(per_sample, beg + chunk);
  const float *yb = y + b * per_sample, *tb = t + b * per_sample, *mb = mask + b * per_sample;
  double yt = 0, yy = 0, tt = 0;
  for (int64_t q = beg + threadIdx.x; q < end; q += blockDim.x) {
    const double m = mb[q], yv = yb[q] * m, tv = tb[q] * m;
    yt += yv * tv;
    yy += yv * yv;
    tt += tv * tv;
  }
  double v[3] = {yt, yy, tt};
  for (int e = 0; e < 3; e++) {
    const double sum = block_sum(v[e], lds);
    if (threadIdx.x == 0) part[((int64_t)b * kCosChunks + ch) * 3 + e] = sum;
  }
}

// one thread per sample: coefficients (g_b, c_b) and the mean loss
__global__ void k_cos_finish(const double *__restrict__ part, int B, double *__restrict__ coef, double *loss) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double tot = 0;
  for (int b = 0; b < B; b++) {
    double yt = 0, yy = 0, tt = 0;
    for (int ch = 0; ch < kCosChunks; ch++) {
      const double *q = part + ((int64_t)b * kCosChunks + ch) * 3;
      yt += q[0];
      yy += q[1];
      tt += q[2];
    }
    double g = 0, cb = 0, lb = 1;
    if (yy > 0 && tt > 0) {
      lb = 1 - yt * yt / (yy * tt);
      g = -2 * yt / (yy * tt * B);
      cb = yt / yy;
    }
    coef[2 * b] = g;
    coef[2 * b + 1] = cb;
    tot += lb;
  }
  *loss = tot / B;
}

__global__ void k_cos_grad(const float *__restrict__ y, const float *__restrict__ t, const float *__restrict__ mask,
                           int64_t per_sample, int B, const double *__restrict__ coef, float *__restrict__ dy) {
  const int64_t n = per_sample * B;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(q / per_sample);
    const double g = coef[2 * b], cb = coef[2 * b + 1];
    dy[q] = (float)(g * ((double)t[q] - cb * (double)y[q]) * (double)mask[q]);
  }
}

// Energy loss per sample b over the crop operator L (the masked 7-point
// Laplacian in cell units, faces closed at the crop edge and next to masked-out
// cells): L_b = 1 - (y, L e)^2 / ((y, L y) (e, L e)), the fraction of the
// error's energy norm that a correction along y, optimally scaled, leaves
// behind (e the target); in [0, 1] by Cauchy-Schwarz.
// dL/dy = -(2 (y,Le) / ((y,Ly)(e,Le) B)) (L e - ((y,Le)/(y,Ly)) L y).
__global__ void k_en_apply(const float *__restrict__ y, const float *__restrict__ mask, int B, int D, int H, int W,
                           float *__restrict__ q) {
  const int64_t n = (int64_t)B * D * H * W;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
    if (mask[c] <= 0) {
      q[c] = 0;
      continue;
    }
    const int x = (int)(c % W), yy = (int)((c / W) % H), z = (int)((c / ((int64_t)W * H)) % D);
    const float yc = y[c];
    float v = 0;
    if (x > 0 && mask[c - 1] > 0) v += yc - y[c - 1];
    if (x < W - 1 && mask[c + 1] > 0) v += yc - y[c + 1];
    if (yy > 0 && mask[c - W] > 0) v += yc - y[c - W];
    if (yy < H - 1 && mask[c + W] > 0) v += yc - y[c + W];
    const int64_t pz = (int64_t)W * H;
    if (z > 0 && mask[c - pz] > 0) v += yc - y[c - pz];
    if (z < D - 1 && mask[c + pz] > 0) v += yc - y[c + pz];
    q[c] = v;
  }
}

// per-(sample, chunk) partials of (y, L e), (y, L y), (e, L e) over the mask
__global__ void k_en_dots(const float *__restrict__ y, const float *__restrict__ qe, const float *__restrict__ q,
                          const float *__restrict__ e, const float *__restrict__ mask, int64_t per_sample,
                          double *__restrict__ part) {
  __shared__ double lds[16];
  const int b = blockIdx.y, ch = blockIdx.x;
  const int64_t chunk = (per_sample + kCosChunks - 1) / kCosChunks;
  const int64_t beg = ch * chunk, end = min(per_sample, beg + chunk);
  const int64_t o = b * per_sample;
  double a = 0, bb = 0, cc = 0;
  for (int64_t t = beg + threadIdx.x; t < end; t += blockDim.x) {
    const int64_t k = o + t;
    if (mask[k] <= 0) continue;
    a += (double)y[k] * qe[k];
    bb += (double)y[k] * q[k];
    cc += (double)e[k] * qe[k];
  }
  double v[3] = {a, bb, cc};
  for (int w = 0; w < 3; w++) {
    const double sum = block_sum(v[w], lds);
    if (threadIdx.x == 0) part[((int64_t)b * kCosChunks + ch) * 3 + w] = sum;
  }
}

__global__ void k_en_finish(const double *__restrict__ part, int B, double *__restrict__ coef, double *loss) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double tot = 0;
  for (int b = 0; b < B; b++) {
    double a = 0, bb = 0, cc = 0;
    for (int ch = 0; ch < kCosChunks; ch++) {
      const double *q = part + ((int64_t)b * kCosChunks + ch) * 3;
      a += q[0];
      bb += q[1];
      cc += q[2];
    }
    double g = 0, cb = 0, lb = 1;
    if (bb > 0 && cc > 0) {
      lb = 1 - a * a / (bb * cc);
      g = -2 * a / (bb * cc * B);
      cb = a / bb;
    }
    coef[2 * b] = g;
    coef[2 * b + 1] = cb;
    tot += lb;
  }
  *loss = tot / B;
}

// qe may alias dy: each thread reads qe[k] before it writes dy[k]
__global__ void k_en_grad(const float *qe, const float *__restrict__ q, const float *__restrict__ mask,
                          int64_t per_sample, int B, const double *__restrict__ coef, float *dy) {
  const int64_t n = per_sample * B;
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(k / per_sample);
    const double v = qe[k];
    dy[k] = mask[k] > 0 ? (float)(coef[2 * b] * (v - coef[2 * b + 1] * (double)q[k])) : 0.f;
  }
}

__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
                       int64_t n, float lr, float b1, float b2, float eps, float c1, float c2) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    const float gq = g[q];
    const float mq = b1 * m[q] + (1 - b1) * gq;
    const float vq = b2 * v[q] + (1 - b2) * gq * gq;
    m[q] = mq;
    v[q] = vq;
    p[q] -= lr * (mq / c1) / (sqrtf(vq / c2) + eps);
  }
}

// ------------------------------------------------------------ the network
struct Level {
  int D, H, W;
};

struct UnetWork {
  int B = 0, D = 0, H = 0, W = 0;
  bf16 *a[9] = {};
  bf16 *p1 = nullptr, *p2 = nullptr;
  float *y = nullptr, *dy = nullptr, *q = nullptr;  // q: L y of the energy loss
  // gradients
  bf16 *g8 = nullptr, *gcat8 = nullptr, *g7 = nullptr, *g6 = nullptr, *gcat6 = nullptr, *g5 = nullptr, *g4 = nullptr,
       *gp2 = nullptr, *g3 = nullptr, *g2 = nullptr, *gp1 = nullptr, *g1 = nullptr, *g0 = nullptr;
  void release() {
    for (auto &p : a) { hipFree(p); p = nullptr; }
    bf16 *gs[] = {p1, p2, g8, gcat8, g7, g6, gcat6, g5, g4, gp2, g3, g2, gp1, g1, g0};
    for (auto p : gs) hipFree(p);
    p1 = p2 = g8 = gcat8 = g7 = g6 = gcat6 = g5 = g4 = gp2 = g3 = g2 = gp1 = g1 = g0 = nullptr;
    hipFree(y); hipFree(dy); hipFree(q);
    y = dy = q = nullptr;
    B = D = H = W = 0;
  }
};

struct xf_unet_s {
  int64_t nparam = 0;
  int64_t off_w[NL], off_b[NL];
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;  // fp32 master, grad, Adam moments
  bf16 *wt[NL] = {};   // forward weights [Cout][Kpad]
  bf16 *wtT[NL] = {};  // input-gradient weights [Cin][Kpad']
  int kpad[NL], kpadT[NL];
  int step = 0;
  int loss_mode = XF_LOSS_MSE;
  double *cos_part = nullptr;  // [B][kCosChunks][3] partial sums of the cosine loss
  int cos_B = 0;
  bool packed = false;
  UnetWork work;
  bool train_buffers = false;
};

typedef struct xf_unet_s *xf_unet;

static int eblocks(int64_t n) {
  int64_t b = (n + 255) / 256;
  return (int)(b > 16384 ? 16384 : (b < 1 ? 1 : b));
}

static void pack(xf_ctx c, xf_unet u) {
  for (int l = 1; l <= 8; l++) {
    hipLaunchKernelGGL(k_pack_weights, dim3(eblocks((int64_t)LCOUT[l] * u->kpad[l])), dim3(256), 0, c->stream,
                       u->p + u->off_w[l], LCOUT[l], LCIN[l], u->wt[l], u->kpad[l]);
    hipLaunchKernelGGL(k_flip_weights, dim3(eblocks((int64_t)LCIN[l] * u->kpadT[l])), dim3(256), 0, c->stream,
                       u->p + u->off_w[l], LCOUT[l], LCIN[l], u->wtT[l], u->kpadT[l]);
  }
  check_kernel("pack weights");
  u->packed = true;
}

// Every wave fetches the block's weight fragments for its own rows, so the
// weight traffic per output cell falls with the rows per wave (TZ TY / 4):
// the widest tile whose input (with its halo) fits the LDS a block may use.
static void tile_of(int cin, int &TZ, int &TY) {
  if (cin <= 64) { TZ = 4; TY = 4; }
  else { TZ = 2; TY = 4; }
}

static size_t conv_lds(int TZ, int TY, int Cin) { return (size_t)(TZ + 2) * (TY + 2) * 18 * (Cin + 8) * sizeof(bf16); }

// LDS a block may use on this device (160 KiB on gfx950)
static size_t conv_lds_max() {
  static const size_t v = [] {
    int dev = 0, bytes = 65536;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&bytes, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
      bytes = 65536;
    return (size_t)std::max(bytes, 65536);
  }();
  return v;
}

// Launch shape: the widest tile (tile_of) and all Cout/16 channel tiles per
// block while that gives at least kConvBlocks blocks; otherwise shorter tiles
// (rows of TZ x TY >= 4, one per wave), then fewer channel tiles per block,
// taking the first shape that reaches kConvBlocks, or the one with most blocks.
// (Each LDS input fragment feeds NT MFMAs: at 32^3, 96 -> 32 channels runs in
// 48 us with NT = 2 on 512 blocks against 65 us with NT = 1 on 1024.)
constexpr int kConvBlocks = 512;
static void conv_shape(int Cin, int Cout, int B, int D, int H, int W, int &TZ, int &TY, int &NT) {
  const int ntmax = Cout / 16;
  int tz0, ty0;
  tile_of(Cin, tz0, ty0);
  const int tiles[4][2] = {{tz0, ty0}, {2, 4}, {1, 4}, {0, 0}};
  int best = -1;
  for (int nt = ntmax; nt >= 1; nt--) {
    if (ntmax % nt) continue;
    for (int q = 0; tiles[q][0]; q++) {
      const int tz = tiles[q][0], ty = tiles[q][1];
      if (q > 0 && tz * ty > tz0 * ty0) continue;
      if (D % tz || H % ty || conv_lds(tz, ty, Cin) > conv_lds_max()) continue;
      const int nb = (W / 16) * (H / ty) * (D / tz) * B * (ntmax / nt);
      if (nb > best) {
        best = nb;
        TZ = tz; TY = ty; NT = nt;
      }
      if (nb >= kConvBlocks) return;
    }
  }
}

// c9: the LDS-tiled kernel where the volume splits into 16 x 4 x 4 tiles
static void conv_last(xf_ctx c, const bf16 *a8, const float *w, const float *bias, const Dst9 &d, int B, int D, int H,
                      int W) {
  if (W % 16 == 0 && H % 4 == 0 && D % 4 == 0)
    hipLaunchKernelGGL(k_conv_last_tile, dim3((W / 16) * (H / 4) * (D / 4) * B), dim3(256), 0, c->stream, a8, w, bias,
                       d, B, D, H, W);
  else
    hipLaunchKernelGGL(k_conv_last, dim3(eblocks((int64_t)B * D * H * W)), dim3(256), 0, c->stream, a8, w, bias, d,
                       B, D, H, W);
}

static void conv(xf_ctx c, const bf16 *src0, int C0, int up0, const bf16 *src1, int C1, const bf16 *wt, int kpad,
                 const float *bias, bf16 *out, int Cout, int B, int D, int H, int W, bool relu) {
  ConvArgs a;
  a.src0 = src0; a.src1 = src1; a.C0 = C0; a.C1 = C1; a.up0 = up0;
  a.wt = wt; a.bias = bias; a.out = out; a.Cout = Cout; a.D = D; a.H = H; a.W = W; a.relu = relu ? 1 : 0;
  a.Kpad = kpad;
  const int Cin = C0 + C1;
  XF_REQUIRE(W % 16 == 0 && H % 4 == 0 && D % 4 == 0 && Cin % 8 == 0 && Cout % 16 == 0 && C0 % 8 == 0, XF_EINVAL,
             "unet conv: unsupported shape");
  XF_REQUIRE(Cin >= 16, XF_EINVAL, "unet conv: at least 16 input channels");
  int NT = Cout / 16;
  a.TZ = a.TY = 0;
  conv_shape(Cin, Cout, B, D, H, W, a.TZ, a.TY, NT);
  a.nsplit = Cout / 16 / NT;
  const size_t lds = conv_lds(a.TZ, a.TY, Cin);
  XF_REQUIRE(lds <= conv_lds_max() && a.TZ * a.TY >= 4 && a.TZ * a.TY <= 16 && (a.TZ * a.TY) % 4 == 0, XF_EINVAL,
             "unet conv: no tile fits in LDS");
  const int nb = (W / 16) * (H / a.TY) * (D / a.TZ) * B * a.nsplit;
  a.m_cpc = fdiv_magic(Cin / 8);
  a.m_sy = fdiv_magic(a.TY + 2);
  const int rpw = a.TZ * a.TY / 4;
#define XF_CONV_CASE(N, R) \
  if (NT == N && rpw == R) { \
    static const bool big_lds_ok = [] { \
      return hipFuncSetAttribute((const void *)k_conv_mfma<N, R>, hipFuncAttributeMaxDynamicSharedMemorySize, \
                                 (int)conv_lds_max()) == hipSuccess; \
    }(); \
    XF_REQUIRE(lds <= 65536 || big_lds_ok, XF_EHIP, "unet conv: the device refused the LDS size"); \
    hipLaunchKernelGGL((k_conv_mfma<N, R>), dim3(nb), dim3(256), lds, c->stream, a); \
    launched = true; \
  }
  bool launched = false;
  XF_CONV_CASE(1, 1) XF_CONV_CASE(1, 2) XF_CONV_CASE(1, 4)
  XF_CONV_CASE(2, 1) XF_CONV_CASE(2, 2) XF_CONV_CASE(2, 4)
  XF_CONV_CASE(3, 1) XF_CONV_CASE(3, 2) XF_CONV_CASE(3, 4)
  XF_CONV_CASE(4, 1) XF_CONV_CASE(4, 2) XF_CONV_CASE(4, 4)
  XF_CONV_CASE(6, 1) XF_CONV_CASE(6, 2) XF_CONV_CASE(6, 4)
#undef XF_CONV_CASE
  XF_REQUIRE(launched, XF_EINVAL, "unet conv: unsupported channel count or tile");
  check_kernel("k_conv_mfma");
}

template <class T>
static T *dalloc(size_t n) {
  T *p = nullptr;
  hipError_t e = hipMalloc(&p, sizeof(T) * (n ? n : 1));
  XF_REQUIRE(e == hipSuccess, XF_ENOMEM, "unet: out of device memory");
  return p;
}

static void ensure_work(xf_unet u, int B, int D, int H, int W, bool train) {
  UnetWork &w = u->work;
  if (w.B == B && w.D == D && w.H == H && w.W == W && (!train || u->train_buffers)) return;
  w.release();
  XF_REQUIRE((int64_t)B * D * H * W < ((int64_t)1 << 31), XF_EINVAL, "unet: more than 2^31 cells in one call");
  XF_REQUIRE(W % 64 == 0 && H % 16 == 0 && D % 16 == 0, XF_EINVAL,
             "unet: the volume must be a multiple of 16 in z, y and of 64 in x (three levels of 16-cell rows)");
  const size_t n1 = (size_t)B * D * H * W, n2 = n1 / 8, n3 = n1 / 64;
  w.B = B; w.D = D; w.H = H; w.W = W;
  w.a[0] = dalloc<bf16>(n1 * 16); w.a[1] = dalloc<bf16>(n1 * 16);
  w.p1 = dalloc<bf16>(n2 * 16);
  w.a[2] = dalloc<bf16>(n2 * 32); w.a[3] = dalloc<bf16>(n2 * 32);
  w.p2 = dalloc<bf16>(n3 * 32);
  w.a[4] = dalloc<bf16>(n3 * 64); w.a[5] = dalloc<bf16>(n3 * 64);
  w.a[6] = dalloc<bf16>(n2 * 32); w.a[7] = dalloc<bf16>(n2 * 32);
  w.a[8] = dalloc<bf16>(n1 * 16);
  w.y = dalloc<float>(n1);
  u->train_buffers = train;
  if (train) {
    w.dy = dalloc<float>(n1);
    w.q = dalloc<float>(n1);
    w.g8 = dalloc<bf16>(n1 * 16); w.gcat8 = dalloc<bf16>(n1 * 48);
    w.g7 = dalloc<bf16>(n2 * 32); w.g6 = dalloc<bf16>(n2 * 32); w.gcat6 = dalloc<bf16>(n2 * 96);
    w.g5 = dalloc<bf16>(n3 * 64); w.g4 = dalloc<bf16>(n3 * 64); w.gp2 = dalloc<bf16>(n3 * 32);
    w.g3 = dalloc<bf16>(n2 * 32); w.g2 = dalloc<bf16>(n2 * 32); w.gp1 = dalloc<bf16>(n2 * 16);
    w.g1 = dalloc<bf16>(n1 * 16); w.g0 = dalloc<bf16>(n1 * 16);
  }
}

// forward through a8; c0's input from `s`
static void forward_body(xf_ctx c, xf_unet u, const Src0 &s, int B) {
  UnetWork &w = u->work;
  const int D = w.D, H = w.H, W = w.W;
  const int64_t n1 = (int64_t)B * D * H * W;
  if (!u->packed) pack(c, u);
  hipLaunchKernelGGL(k_conv_first, dim3((int)(n1 / 256)), dim3(256), 0, c->stream, s, u->p + u->off_w[0],
                     u->p + u->off_b[0], w.a[0], B);
  check_kernel("k_conv_first");
  const float *P = u->p;
  conv(c, w.a[0], 16, 0, nullptr, 0, u->wt[1], u->kpad[1], P + u->off_b[1], w.a[1], 16, B, D, H, W, true);
  hipLaunchKernelGGL(k_pool, dim3(eblocks(n1 / 8 * 16)), dim3(256), 0, c->stream, w.a[1], w.p1, B, D, H, W, 16);
  conv(c, w.p1, 16, 0, nullptr, 0, u->wt[2], u->kpad[2], P + u->off_b[2], w.a[2], 32, B, D / 2, H / 2, W / 2, true);
  conv(c, w.a[2], 32, 0, nullptr, 0, u->wt[3], u->kpad[3], P + u->off_b[3], w.a[3], 32, B, D / 2, H / 2, W / 2, true);
  hipLaunchKernelGGL(k_pool, dim3(eblocks(n1 / 64 * 32)), dim3(256), 0, c->stream, w.a[3], w.p2, B, D / 2, H / 2,
                     W / 2, 32);
  conv(c, w.p2, 32, 0, nullptr, 0, u->wt[4], u->kpad[4], P + u->off_b[4], w.a[4], 64, B, D / 4, H / 4, W / 4, true);
  conv(c, w.a[4], 64, 0, nullptr, 0, u->wt[5], u->kpad[5], P + u->off_b[5], w.a[5], 64, B, D / 4, H / 4, W / 4, true);
  conv(c, w.a[5], 64, 1, w.a[3], 32, u->wt[6], u->kpad[6], P + u->off_b[6], w.a[6], 32, B, D / 2, H / 2, W / 2, true);
  conv(c, w.a[6], 32, 0, nullptr, 0, u->wt[7], u->kpad[7], P + u->off_b[7], w.a[7], 32, B, D / 2, H / 2, W / 2, true);
  conv(c, w.a[7], 32, 1, w.a[1], 16, u->wt[8], u->kpad[8], P + u->off_b[8], w.a[8], 16, B, D, H, W, true);
}

// ------------------------------------------------------------------ C ABI
extern "C" int xf_unet_create(xf_ctx c, int seed, xf_unet *out) {
  xf_unet u = nullptr;
  int st = guarded(c, [&] {
    XF_REQUIRE(out, XF_EINVAL, "xf_unet_create: null out");
    u = new xf_unet_s();
    int64_t off = 0;
    for (int l = 0; l < NL; l++) {
      u->off_w[l] = off;
      off += (int64_t)LCOUT[l] * 27 * LCIN[l];
      u->off_b[l] = off;
      off += LCOUT[l];
      u->kpad[l] = kpad_of(LCIN[l]);
      u->kpadT[l] = kpad_of(LCOUT[l]);
    }
    u->nparam = off;
    // He-normal initialisation (fan_in = 27 Cin), zero biases; the last layer
    // starts small so the untrained network is a mild smoother
    std::mt19937_64 rng(seed);
    std::vector<float> h(off, 0.0f);
    for (int l = 0; l < NL; l++) {
      std::normal_distribution<float> nd(0.0f, std::sqrt(2.0f / (27.0f * LCIN[l])) * (l == NL - 1 ? 0.1f : 1.0f));
      for (int64_t q = 0; q < (int64_t)LCOUT[l] * 27 * LCIN[l]; q++) h[u->off_w[l] + q] = nd(rng);
    }
    u->p = dalloc<float>(off); u->g = dalloc<float>(off); u->m = dalloc<float>(off); u->v = dalloc<float>(off);
    XF_HIP(hipMemcpy(u->p, h.data(), sizeof(float) * off, hipMemcpyHostToDevice));
    XF_HIP(hipMemset(u->m, 0, sizeof(float) * off));
    XF_HIP(hipMemset(u->v, 0, sizeof(float) * off));
    for (int l = 1; l <= 8; l++) {
      u->wt[l] = dalloc<bf16>((size_t)LCOUT[l] * u->kpad[l]);
      u->wtT[l] = dalloc<bf16>((size_t)LCIN[l] * u->kpadT[l]);
    }
  });
  if (st != XF_OK) {
    delete u;
    return st;
  }
  *out = u;
  return XF_OK;
}

extern "C" int xf_unet_destroy(xf_ctx c, xf_unet u) {
  return guarded(c, [&] {
    if (!u) return;
    XF_HIP(hipStreamSynchronize(c->stream));
    u->work.release();
    hipFree(u->p); hipFree(u->g); hipFree(u->m); hipFree(u->v); hipFree(u->cos_part);
    for (int l = 0; l < NL; l++) { hipFree(u->wt[l]); hipFree(u->wtT[l]); }
    delete u;
  });
}

extern "C" int xf_unet_param_count(xf_unet u, int64_t *n) {
  if (!u || !n) return XF_EINVAL;
  *n = u->nparam;
  return XF_OK;
}

extern "C" int xf_unet_get_params(xf_ctx c, xf_unet u, float *host) {
  return guarded(c, [&] {
    XF_HIP(hipMemcpyAsync(host, u->p, sizeof(float) * u->nparam, hipMemcpyDeviceToHost, c->stream));
    XF_HIP(hipStreamSynchronize(c->stream));
  });
}

extern "C" int xf_unet_set_params(xf_ctx c, xf_unet u, const float *host) {
  return guarded(c, [&] {
    XF_HIP(hipMemcpyAsync(u->p, host, sizeof(float) * u->nparam, hipMemcpyHostToDevice, c->stream));
    u->packed = false;
    XF_HIP(hipStreamSynchronize(c->stream));
  });
}

extern "C" int xf_unet_get_grads(xf_ctx c, xf_unet u, float *host) {
  return guarded(c, [&] {
    XF_HIP(hipMemcpyAsync(host, u->g, sizeof(float) * u->nparam, hipMemcpyDeviceToHost, c->stream));
    XF_HIP(hipStreamSynchronize(c->stream));
  });
}

// Crop forward: r, mask [B][D][H][W] fp32 (device); y [B][D][H][W] fp32.
extern "C" int xf_unet_forward(xf_ctx c, xf_unet u, int B, int D, int H, int W, const float *r, const float *mask,
                               float *y) {
  return guarded(c, [&] {
    XF_REQUIRE(u && r && mask && y && B > 0, XF_EINVAL, "xf_unet_forward: bad argument");
    ensure_work(u, B, D, H, W, false);
    Src0 s{r, mask, nullptr, nullptr, 0, 0, D, H, W};
    forward_body(c, u, s, B);
    Dst9 d{y, nullptr, nullptr, 0, 0};
    conv_last(c, u->work.a[8], u->p + u->off_w[9], u->p + u->off_b[9], d, B, D, H, W);
    check_kernel("k_conv_last");
  });
}

static void wgrad(xf_ctx c, xf_unet u, int l, const bf16 *src0, int C0, int up0, const bf16 *src1, int C1,
                  const bf16 *g, const float *g32, int B, int D, int H, int W) {
  GradArgs a;
  a.src0 = src0; a.src1 = src1; a.C0 = C0; a.C1 = C1; a.up0 = up0;
  a.g = g; a.g32 = g32; a.Cout = LCOUT[l]; a.D = D; a.H = H; a.W = W; a.B = B;
  a.dw = u->g + u->off_w[l];
  a.db = u->g + u->off_b[l];
  const int64_t ncell = (int64_t)B * D * H * W;
  int chunks = (int)std::max<int64_t>(1, std::min<int64_t>(64, ncell / 4096));
  hipLaunchKernelGGL(k_conv_wgrad, dim3(chunks, LCOUT[l], 27 * ((C0 + C1) / 8)), dim3(256), 0, c->stream, a, chunks);
  check_kernel("k_conv_wgrad");
}

// First-layer weight gradient: the input is the 2-channel (r, mask) crop.
__global__ void k_first_wgrad(Src0 s, const bf16 *__restrict__ g0, int B, float *dw, float *db, int chunks) {
  __shared__ float red[256];
  const int co = blockIdx.y;
  const int64_t ncell = (int64_t)B * s.D * s.H * s.W;
  const int64_t per = (ncell + chunks - 1) / chunks;
  const int64_t beg = blockIdx.x * per, end = min(ncell, beg + per);
  const float inv = src0_inv(s);
  float acc[54];
  for (int i = 0; i < 54; i++) acc[i] = 0;
  float gb = 0;
  for (int64_t q = beg + threadIdx.x; q < end; q += blockDim.x) {
    const float gv = from_bf16(g0[q * 16 + co]);
    if (gv == 0.0f) continue;
    const int x = (int)(q % s.W), y = (int)((q / s.W) % s.H), z = (int)((q / ((int64_t)s.W * s.H)) % s.D);
    const int b = (int)(q / ((int64_t)s.W * s.H * s.D));
    gb += gv;
    for (int t = 0; t < 27; t++) {
      float rv, mv;
      src0_read(s, b, z + t / 9 - 1, y + (t / 3) % 3 - 1, x + t % 3 - 1, inv, rv, mv);
      acc[2 * t] += gv * rv;
      acc[2 * t + 1] += gv * mv;
    }
  }
  for (int i = 0; i <= 54; i++) {
    red[threadIdx.x] = i < 54 ? acc[i] : gb;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
      if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      if (i < 54) atomicAdd(&dw[co * 54 + i], red[0]);
      else atomicAdd(&db[co], red[0]);
    }
    __syncthreads();
  }
}

// Loss and gradients of a crop batch (no update): mean over mask of (y - t)^2.
static double backward(xf_ctx c, xf_unet u, int B, int D, int H, int W, const float *r, const float *mask,
                       const float *target) {
  ensure_work(u, B, D, H, W, true);
  UnetWork &w = u->work;
  Src0 s{r, mask, nullptr, nullptr, 0, 0, D, H, W};
  forward_body(c, u, s, B);
  const int64_t n1 = (int64_t)B * D * H * W;
  Dst9 d{w.y, nullptr, nullptr, 0, 0};
  conv_last(c, w.a[8], u->p + u->off_w[9], u->p + u->off_b[9], d, B, D, H, W);
  if (u->loss_mode == XF_LOSS_COSINE || u->loss_mode == XF_LOSS_ENERGY) {
    if (u->cos_B < B) {
      hipFree(u->cos_part);
      u->cos_part = nullptr;
      u->cos_part = dalloc<double>((size_t)B * kCosChunks * 3 + 2 * (size_t)B);
      u->cos_B = B;
    }
    const int64_t per = (int64_t)D * H * W;
    double *coef = u->cos_part + (size_t)B * kCosChunks * 3;
    if (u->loss_mode == XF_LOSS_ENERGY) {
      // q = L y into w.q, L e into w.dy (overwritten in place by the gradient)
      hipLaunchKernelGGL(k_en_apply, dim3(std::min(eblocks(n1), 4096)), dim3(256), 0, c->stream, w.y, mask, B, D, H, W,
                         w.q);
      hipLaunchKernelGGL(k_en_apply, dim3(std::min(eblocks(n1), 4096)), dim3(256), 0, c->stream, target, mask, B, D,
                         H, W, w.dy);
      hipLaunchKernelGGL(k_en_dots, dim3(kCosChunks, B), dim3(256), 0, c->stream, w.y, w.dy, w.q, target, mask, per,
                         u->cos_part);
      hipLaunchKernelGGL(k_en_finish, dim3(1), dim3(64), 0, c->stream, u->cos_part, B, coef, c->d_scalar);
      hipLaunchKernelGGL(k_en_grad, dim3(std::min(eblocks(n1), 4096)), dim3(256), 0, c->stream, w.dy, w.q, mask, per,
                         B, coef, w.dy);
    } else {
    hipLaunchKernelGGL(k_cos_dots, dim3(kCosChunks, B), dim3(256), 0, c->stream, w.y, target, mask, per, u->cos_part);
    hipLaunchKernelGGL(k_cos_finish, dim3(1), dim3(64), 0, c->stream, u->cos_part, B, coef, c->d_scalar);
    hipLaunchKernelGGL(k_cos_grad, dim3(std::min(eblocks(n1), 4096)), dim3(256), 0, c->stream, w.y, target, mask, per,
                       B, coef, w.dy);
    }
  } else {
    // mean over the masked cells: the mask sum first, then the loss
    hipLaunchKernelGGL(k_sum_f32, dim3(std::min(eblocks(n1), 2048)), dim3(256), 0, c->stream, mask, n1, c->partials,
                       c->counters, c->d_scalar + 1);
    hipLaunchKernelGGL(k_mse, dim3(std::min(eblocks(n1), 2048)), dim3(256), 0, c->stream, w.y, target, mask, n1, w.dy,
                       c->partials + kMaxPartials, c->counters + 1, c->d_scalar, c->d_scalar + 1);
  }
  XF_HIP(hipMemsetAsync(u->g, 0, sizeof(float) * u->nparam, c->stream));
  const int D2 = D / 2, H2 = H / 2, W2 = W / 2, D4 = D / 4, H4 = H / 4, W4 = W / 4;
  const int64_t n2 = n1 / 8, n3 = n1 / 64;
  // c9
  wgrad(c, u, 9, w.a[8], 16, 0, nullptr, 0, nullptr, w.dy, B, D, H, W);
  hipLaunchKernelGGL(k_last_dx, dim3(eblocks(n1)), dim3(256), 0, c->stream, w.dy, u->p + u->off_w[9], w.g8, B, D, H, W);
  // c8: input [up a7 (32), a1 (16)]
  hipLaunchKernelGGL(k_relu_back, dim3(eblocks(n1 * 16)), dim3(256), 0, c->stream, w.g8, w.a[8], n1 * 16);
  wgrad(c, u, 8, w.a[7], 32, 1, w.a[1], 16, w.g8, nullptr, B, D, H, W);
  conv(c, w.g8, 16, 0, nullptr, 0, u->wtT[8], u->kpadT[8], nullptr, w.gcat8, 48, B, D, H, W, false);
  hipLaunchKernelGGL(k_upsample_back, dim3(eblocks(n2 * 32)), dim3(256), 0, c->stream, w.gcat8, 48, 0, w.g7, 32, B,
                     D2, H2, W2);
  // c7
  hipLaunchKernelGGL(k_relu_back, dim3(eblocks(n2 * 32)), dim3(256), 0, c->stream, w.g7, w.a[7], n2 * 32);
  wgrad(c, u, 7, w.a[6], 32, 0, nullptr, 0, w.g7, nullptr, B, D2, H2, W2);
  conv(c, w.g7, 32, 0, nullptr, 0, u->wtT[7], u->kpadT[7], nullptr, w.g6, 32, B, D2, H2, W2, false);
  // c6: input [up a5 (64), a3 (32)]
  hipLaunchKernelGGL(k_relu_back, dim3(eblocks(n2 * 32)), dim3(256), 0, c->stream, w.g6, w.a[6], n2 * 32);
  wgrad(c, u, 6, w.a[5], 64, 1, w.a[3], 32, w.g6, nullptr, B, D2, H2, W2);
  conv(c, w.g6, 32, 0, nullptr, 0, u->wtT[6], u->kpadT[6], nullptr, w.gcat6, 96, B, D2, H2, W2, false);
  hipLaunchKernelGGL(k_upsample_back, dim3(eblocks(n3 * 64)), dim3(256), 0, c->stream, w.gcat6, 96, 0, w.g5, 64, B,
                     D4, H4, W4);
  // c5
  hipLaunchKernelGGL(k_relu_back, dim3(eblocks(n3 * 64)), dim3(256), 0, c->stream, w.g5, w.a[5], n3 * 64);
  wgrad(c, u, 5, w.a[4], 64, 0, nullptr, 0, w.g5, nullptr, B, D4, H4, W4);
  conv(c, w.g5, 64, 0, nullptr, 0, u->wtT[5], u->kpadT[5], nullptr, w.g4, 64, B, D4, H4, W4, false);
  // c4 (input p2 = pool a3)
  hipLaunchKernelGGL(k_relu_back, dim3(eblocks(n3 * 64)), dim3(256), 0, c->stream, w.g4, w.a[4], n3 * 64);
  wgrad(c, u, 4, w.p2, 32, 0, nullptr, 0, w.g4, nullptr, B, D4, H4, W4);
  conv(c, w.g4, 64, 0, nullptr, 0, u->wtT[4], u->kpadT[4], nullptr, w.gp2, 32, B, D4, H4, W4, false);
  // a3 gradient: concat slice of c6 + pool backward of gp2
  hipLaunchKernelGGL(k_skip_grad, dim3(eblocks(n2 * 32)), dim3(256), 0, c->stream, w.gcat6, 96, 64, w.gp2, w.g3, 32, B,
                     D2, H2, W2);
  // c3
  hipLaunchKernelGGL(k_relu_back, dim3(eblocks(n2 * 32)), dim3(256), 0, c->stream, w.g3, w.a[3], n2 * 32);
  wgrad(c, u, 3, w.a[2], 32, 0, nullptr, 0, w.g3, nullptr, B, D2, H2, W2);
  conv(c, w.g3, 32, 0, nullptr, 0, u->wtT[3], u->kpadT[3], nullptr, w.g2, 32, B, D2, H2, W2, false);
  // c2 (input p1 = pool a1)
  hipLaunchKernelGGL(k_relu_back, dim3(eblocks(n2 * 32)), dim3(256), 0, c->stream, w.g2, w.a[2], n2 * 32);
  wgrad(c, u, 2, w.p1, 16, 0, nullptr, 0, w.g2, nullptr, B, D2, H2, W2);
  conv(c, w.g2, 32, 0, nullptr, 0, u->wtT[2], u->kpadT[2], nullptr, w.gp1, 16, B, D2, H2, W2, false);
  // a1 gradient: concat slice of c8 + pool backward of gp1
  hipLaunchKernelGGL(k_skip_grad, dim3(eblocks(n1 * 16)), dim3(256), 0, c->stream, w.gcat8, 48, 32, w.gp1, w.g1, 16, B,
                     D, H, W);
  // c1
  hipLaunchKernelGGL(k_relu_back, dim3(eblocks(n1 * 16)), dim3(256), 0, c->stream, w.g1, w.a[1], n1 * 16);
  wgrad(c, u, 1, w.a[0], 16, 0, nullptr, 0, w.g1, nullptr, B, D, H, W);
  conv(c, w.g1, 16, 0, nullptr, 0, u->wtT[1], u->kpadT[1], nullptr, w.g0, 16, B, D, H, W, false);
  // c0
  hipLaunchKernelGGL(k_relu_back, dim3(eblocks(n1 * 16)), dim3(256), 0, c->stream, w.g0, w.a[0], n1 * 16);
  {
    int chunks = (int)std::max<int64_t>(1, std::min<int64_t>(64, n1 / 4096));
    hipLaunchKernelGGL(k_first_wgrad, dim3(chunks, 16), dim3(256), 0, c->stream, s, w.g0, B, u->g + u->off_w[0],
                       u->g + u->off_b[0], chunks);
  }
  check_kernel("unet backward");
  double loss = 0;
  XF_HIP(hipMemcpyAsync(&loss, c->d_scalar, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  XF_HIP(hipStreamSynchronize(c->stream));
  return loss;
}

extern "C" int xf_unet_set_loss(xf_ctx c, xf_unet u, int mode) {
  return guarded(c, [&] {
    XF_REQUIRE(u && (mode == XF_LOSS_MSE || mode == XF_LOSS_COSINE || mode == XF_LOSS_ENERGY), XF_EINVAL,
               "xf_unet_set_loss: bad mode");
    u->loss_mode = mode;
  });
}

extern "C" int xf_unet_loss_grad(xf_ctx c, xf_unet u, int B, int D, int H, int W, const float *r, const float *mask,
                                 const float *target, double *loss) {
  return guarded(c, [&] {
    XF_REQUIRE(u && r && mask && target && B > 0, XF_EINVAL, "xf_unet_loss_grad: bad argument");
    double l = backward(c, u, B, D, H, W, r, mask, target);
    if (loss) *loss = l;
  });
}

extern "C" int xf_unet_train_step(xf_ctx c, xf_unet u, int B, int D, int H, int W, const float *r, const float *mask,
                                  const float *target, float lr, double *loss) {
  return guarded(c, [&] {
    XF_REQUIRE(u && r && mask && target && B > 0 && lr > 0, XF_EINVAL, "xf_unet_train_step: bad argument");
    double l = backward(c, u, B, D, H, W, r, mask, target);
    u->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - powf(b1, (float)u->step), c2 = 1.0f - powf(b2, (float)u->step);
    hipLaunchKernelGGL(k_adam, dim3(eblocks(u->nparam)), dim3(256), 0, c->stream, u->p, u->g, u->m, u->v, u->nparam,
                       lr, b1, b2, 1e-8f, c1, c2);
    check_kernel("k_adam");
    u->packed = false;
    if (loss) *loss = l;
  });
}

// Preconditioner application on a pressure grid: z = s * net(r / s) on the
// unknowns (0 elsewhere), s^2 = mean of r^2 over the unknowns.
void unet_apply_grid(xf_ctx c, xf_unet u, xf_grid g, const float *r32, float *z32, const double *d_sumsq) {
  const int D = g->n - 2, H = g->l - 2, W = g->m - 2;
  ensure_work(u, 1, D, H, W, false);
  Src0 s{r32, nullptr, g->d_flags, d_sumsq, (double)g->nunk, 1, D, H, W};
  forward_body(c, u, s, 1);
  Dst9 d{z32, g->d_flags, d_sumsq, (double)g->nunk, 1};
  conv_last(c, u->work.a[8], u->p + u->off_w[9], u->p + u->off_b[9], d, 1, D, H, W);
  check_kernel("k_conv_last");
}

__global__ void k_sumsq_unknowns_f32(int64_t nc, const uint8_t *__restrict__ flags, const float *__restrict__ v,
                                     double *partials, unsigned *counter, double *out) {
  __shared__ double lds[16];
  double acc = 0;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x)
    if (flags[c] & FL_UNKNOWN) acc += (double)v[c] * (double)v[c];
  double bs = block_sum(acc, lds);
  double tot;
  if (grid_sum_last(bs, partials, counter, &tot, lds) && threadIdx.x == 0) *out = tot;
}

extern "C" int xf_unet_apply(xf_ctx c, xf_unet u, xf_grid g, const float *r, float *z) {
  return guarded(c, [&] {
    XF_REQUIRE(u && g && r && z, XF_EINVAL, "xf_unet_apply: null argument");
    XF_REQUIRE(r != z, XF_EINVAL, "xf_unet_apply: z must not alias r");
    hipLaunchKernelGGL(k_sumsq_unknowns_f32, dim3(std::min(eblocks(g->ncell), 2048)), dim3(256), 0, c->stream,
                       g->ncell, g->d_flags, r, c->partials, c->counters, c->d_scalar + 2);
    unet_apply_grid(c, u, g, r, z, c->d_scalar + 2);
  });
}
