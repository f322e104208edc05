// Learned warm-start network: 3-D conv U-Net training and inference on
// CDNA4 bf16 MFMA, fp32 accumulation: the convolutions step K by 32 with
// v_mfma_f32_16x16x32_bf16 (gfx950) and finish a K that is 16 short of a
// multiple of 32 with one v_mfma_f32_16x16x16_bf16.
//
// Layout: activations channels-last, cell index x + nx (y + ny (z + nz b)),
// bf16; activation gradients fp32.  A 3x3x3 convolution is an implicit GEMM
// with M = cells, N = output channels, K = 27 taps x input channels (+ the
// bias column, whose input is 1): each wave computes 16 cells x 16 channels
// per MFMA tile.  The input gradient is the same GEMM with the weights
// transposed and the taps mirrored; the weight gradient is a GEMM over cells
// (K = cells) written as per-chunk partials and summed in a fixed order, so
// training is deterministic.  No counterpart in the reference.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>
#include <math.h>
#include <algorithm>
#include <string>
#include <vector>
#include "../../include/waves_net.h"
#include "wv_netdev.h"

typedef uint16_t bf16_t;
typedef short v4s __attribute__((ext_vector_type(4)));
typedef short v8s __attribute__((ext_vector_type(8)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));
// 16x16 output tile += A (16 x 32) B (32 x 16): lane l holds A[l % 16][8 (l / 16) + 0..7]
// and B[8 (l / 16) + 0..7][l % 16]; the accumulator layout is the 16x16x16 one
__device__ __forceinline__ v4f mfma32(v8s a, v8s b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, a), __builtin_bit_cast(v8bf, b), c, 0, 0, 0);
}

struct Grid { int nx, ny, nz, B; };

static inline int64_t cells_of(Grid g) { return (int64_t)g.nx * g.ny * g.nz * g.B; }
static inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

__device__ __forceinline__ float bf2f(bf16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ bf16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (bf16_t)0x7fc0;
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16_t)(u >> 16);
}
__device__ __forceinline__ int64_t gcells(const Grid& g) { return (int64_t)g.nx * g.ny * g.nz * g.B; }
__device__ __forceinline__ void decode(int64_t c, const Grid& g, int& x, int& y, int& z, int& b) {
  x = (int)(c % g.nx); int64_t t = c / g.nx;
  y = (int)(t % g.ny); t /= g.ny;
  z = (int)(t % g.nz); b = (int)(t / g.nz);
}
__device__ __forceinline__ int64_t encode(const Grid& g, int x, int y, int z, int b) {
  return x + (int64_t)g.nx * (y + (int64_t)g.ny * (z + (int64_t)g.nz * b));
}
__device__ __forceinline__ bool inside(const Grid& g, int x, int y, int z) {
  return x >= 0 && x < g.nx && y >= 0 && y < g.ny && z >= 0 && z < g.nz;
}

// ------------------------------------------------------------ convolution --
struct ConvArgs {
  const bf16_t* xa; const bf16_t* xb; const bf16_t* w;
  bf16_t* yb; float* yf;
  Grid g, gb;
  int ca, cb, kp, cout, ystride, bias, relu;
  int groups;   // k_conv_tiled: output-channel groups of NT x 16 channels, one block each (0 or 1: one group)
};

// y[c][co] = act(sum_k A[c][k] W[co][k]); A[c][t cin + ci] = input ci at the
// tap-t neighbour of c (zero outside), A[c][27 cin] = bias input.
template <int NT>
__global__ void __launch_bounds__(256) k_conv(ConvArgs a) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m = lane & 15, kq = lane >> 4;
  const int64_t M = gcells(a.g);
  const int64_t cell0 = ((int64_t)blockIdx.x * 4 + wave) * 16;
  if (cell0 >= M) return;                      // wave-uniform; no barriers below
  const int cin = a.ca + a.cb, kreal = 27 * cin;
  const int64_t c = cell0 + m;
  const bool live = c < M;
  int x = 0, y = 0, z = 0, b = 0;
  if (live) decode(c, a.g, x, y, z, b);
  v4f acc[NT];
  for (int t = 0; t < NT; t++) acc[t] = v4f{0.f, 0.f, 0.f, 0.f};
  // the A fragment of KW consecutive k from k (within one tap: cin, ca and cb
  // are multiples of 8), the bias column's 1 at k == kreal
  auto a_frag = [&](int k, auto zero) {
    using V = decltype(zero);
    V av = zero;
    if (live && k < kreal) {
      const int tap = k / cin, ch = k - tap * cin;
      const int qx = x + tap % 3 - 1, qy = y + (tap / 3) % 3 - 1, qz = z + tap / 9 - 1;
      if (inside(a.g, qx, qy, qz)) {
        const bf16_t* src = ch < a.ca ? a.xa + encode(a.g, qx, qy, qz, b) * a.ca + ch
                                      : a.xb + encode(a.gb, qx >> 1, qy >> 1, qz >> 1, b) * a.cb + (ch - a.ca);
        av = *(const V*)src;
      }
    } else if (live && k == kreal && a.bias) {
      av[0] = (short)0x3f80;
    }
    return av;
  };
  int k0 = 0;
  for (; k0 + 32 <= a.kp; k0 += 32) {
    const int k = k0 + 8 * kq;
    const v8s av = a_frag(k, v8s{0, 0, 0, 0, 0, 0, 0, 0});
    for (int t = 0; t < NT; t++) {
      const int co = t * 16 + m;
      v8s bv = {0, 0, 0, 0, 0, 0, 0, 0};
      if (co < a.cout) bv = *(const v8s*)(a.w + (int64_t)co * a.kp + k);
      acc[t] = mfma32(av, bv, acc[t]);
    }
  }
  if (k0 < a.kp) {                             // kp = 32 j + 16
    const int k = k0 + 4 * kq;
    const v4s av = a_frag(k, v4s{0, 0, 0, 0});
    for (int t = 0; t < NT; t++) {
      const int co = t * 16 + m;
      v4s bv = {0, 0, 0, 0};
      if (co < a.cout) bv = *(const v4s*)(a.w + (int64_t)co * a.kp + k);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[t], 0, 0, 0);
    }
  }
  for (int t = 0; t < NT; t++)
    for (int j = 0; j < 4; j++) {
      const int64_t cc = cell0 + 4 * kq + j;
      const int co = t * 16 + m;
      if (cc < M && co < a.cout) {
        float v = acc[t][j];
        if (a.relu) v = v > 0.f ? v : 0.f;
        if (a.yb) a.yb[cc * a.ystride + co] = f2bf(v);
        else a.yf[cc * a.ystride + co] = v;
      }
    }
}

// Forward convolution with the input halo staged in LDS: a block computes a
// 16 x 4 x 4 tile of output cells; its 18 x 6 x 6 input halo (every input
// channel, the upsampled coarse part included) is read from HBM once and the
// 27 taps read it from LDS.  Wave w computes the rows (y, z) = (w, 0..3) of
// the tile as four 16-cell MFMA tiles sharing each weight fragment.
#define TX 16
#define TY 4
#define TZ 4
#define HX (TX + 2)
#define HY (TY + 2)
#define HZ (TZ + 2)
#define HALO (HX * HY * HZ)
template <int CIN, int NT, bool WLDS>
__global__ void __launch_bounds__(256) k_conv_tiled(ConvArgs a) {
  extern __shared__ __attribute__((aligned(16))) bf16_t lds[];
  bf16_t* halo = lds;                                   // HALO * CIN
  bf16_t* wl = lds + HALO * CIN;                        // NT*16 rows x kp (WLDS)
  const int tilesz = (a.g.nz + TZ - 1) / TZ, groups = a.groups > 1 ? a.groups : 1;
  const int zb = (int)blockIdx.z / groups, co0 = ((int)blockIdx.z - zb * groups) * NT * 16;   // this block's channels
  const int x0 = blockIdx.x * TX, y0 = blockIdx.y * TY;
  const int z0 = (zb % tilesz) * TZ, b = zb / tilesz;
  constexpr int Q = CIN / 4;                            // 8-byte chunks per halo cell
  constexpr int TOTAL = HALO * Q, PER = 8;              // PER loads in flight per thread
  for (int base = 0; base < TOTAL; base += 256 * PER) {
    v4s v[PER];
#pragma unroll
    for (int u = 0; u < PER; u++) {
      const int e = base + u * 256 + threadIdx.x;
      v[u] = v4s{0, 0, 0, 0};
      if (e < TOTAL) {
        const int hc = e / Q, ch = (e - hc * Q) * 4;
        const int hx = hc % HX, hy = (hc / HX) % HY, hz = hc / (HX * HY);
        const int x = x0 - 1 + hx, y = y0 - 1 + hy, z = z0 - 1 + hz;
        if (inside(a.g, x, y, z))
          v[u] = ch < a.ca ? *(const v4s*)(a.xa + encode(a.g, x, y, z, b) * a.ca + ch)
                           : *(const v4s*)(a.xb + encode(a.gb, x >> 1, y >> 1, z >> 1, b) * a.cb + (ch - a.ca));
      }
    }
#pragma unroll
    for (int u = 0; u < PER; u++) {
      const int e = base + u * 256 + threadIdx.x;
      if (e < TOTAL) *(v4s*)(halo + e * 4) = v[u];       // hc * CIN + ch == 4 e
    }
  }
  if (WLDS) {
    const int kq4 = a.kp / 4, wtotal = NT * 16 * kq4;
    for (int base = 0; base < wtotal; base += 256 * PER) {
      v4s v[PER];
#pragma unroll
      for (int u = 0; u < PER; u++) {
        const int e = base + u * 256 + threadIdx.x;
        const int row = e / kq4, k = (e - row * kq4) * 4;
        v[u] = (e < wtotal && co0 + row < a.cout) ? *(const v4s*)(a.w + (int64_t)(co0 + row) * a.kp + k) : v4s{0, 0, 0, 0};
      }
#pragma unroll
      for (int u = 0; u < PER; u++) {
        const int e = base + u * 256 + threadIdx.x;
        if (e < wtotal) *(v4s*)(wl + e * 4) = v[u];         // row * kp + k == 4 e
      }
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m = lane & 15, kq = lane >> 4;
  constexpr int KREAL = 27 * CIN;
  v4f acc[4][NT];
  for (int r = 0; r < 4; r++) for (int t = 0; t < NT; t++) acc[r][t] = v4f{0.f, 0.f, 0.f, 0.f};
  static_assert(CIN % 8 == 0, "an 8-wide K fragment stays within one tap");
  // halo offset of this lane's fragment at k, or -1 past the taps
  auto tap_offset = [&](int k) {
    if (k >= KREAL) return -1;
    const int tap = k / CIN, ch = k - tap * CIN;
    return ((tap / 9) * (HX * HY) + ((tap / 3) % 3) * HX + (tap % 3) + m) * CIN + ch;
  };
  int k0 = 0;
#pragma unroll 2
  for (; k0 + 32 <= a.kp; k0 += 32) {
    const int k = k0 + 8 * kq;
    v8s bv[NT];
    for (int t = 0; t < NT; t++) {
      const int cl = t * 16 + m, co = co0 + cl;
      if (WLDS) bv[t] = *(const v8s*)(wl + cl * a.kp + k);
      else bv[t] = co < a.cout ? *(const v8s*)(a.w + (int64_t)co * a.kp + k) : v8s{0, 0, 0, 0, 0, 0, 0, 0};
    }
    const int off = tap_offset(k);
    for (int r = 0; r < 4; r++) {                       // row r of this wave: y = wave, z = r
      v8s av = {0, 0, 0, 0, 0, 0, 0, 0};
      if (off >= 0) av = *(const v8s*)(halo + off + (r * (HX * HY) + wave * HX) * CIN);
      else if (k == KREAL) av[0] = (short)0x3f80;
      for (int t = 0; t < NT; t++) acc[r][t] = mfma32(av, bv[t], acc[r][t]);
    }
  }
  if (k0 < a.kp) {                                     // kp = 32 j + 16
    const int k = k0 + 4 * kq;
    v4s bv[NT];
    for (int t = 0; t < NT; t++) {
      const int cl = t * 16 + m, co = co0 + cl;
      if (WLDS) bv[t] = *(const v4s*)(wl + cl * a.kp + k);
      else bv[t] = co < a.cout ? *(const v4s*)(a.w + (int64_t)co * a.kp + k) : v4s{0, 0, 0, 0};
    }
    const int off = tap_offset(k);
    for (int r = 0; r < 4; r++) {
      v4s av = {0, 0, 0, 0};
      if (off >= 0) av = *(const v4s*)(halo + off + (r * (HX * HY) + wave * HX) * CIN);
      else if (k == KREAL) av = v4s{(short)0x3f80, 0, 0, 0};
      for (int t = 0; t < NT; t++) acc[r][t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv[t], acc[r][t], 0, 0, 0);
    }
  }
  const int y = y0 + wave;
  for (int r = 0; r < 4; r++) {
    const int z = z0 + r;
    if (y >= a.g.ny || z >= a.g.nz) continue;
    for (int t = 0; t < NT; t++)
      for (int j = 0; j < 4; j++) {
        const int x = x0 + 4 * kq + j, co = co0 + t * 16 + m;
        if (x < a.g.nx && co < a.cout) {
          float v = acc[r][t][j];
          if (a.relu) v = v > 0.f ? v : 0.f;
          const int64_t cc = encode(a.g, x, y, z, b);
          if (a.yb) a.yb[cc * a.ystride + co] = f2bf(v);
          else a.yf[cc * a.ystride + co] = v;
        }
      }
  }
}

static int lds_limit() {
  static int limit = -1;
  if (limit < 0) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) v = 65536;
    limit = v;
  }
  return limit;
}

template <int CIN, int NT>
static bool launch_tiled_nt(const ConvArgs& a, hipStream_t st) {
  const dim3 grid((a.g.nx + TX - 1) / TX, (a.g.ny + TY - 1) / TY, ((a.g.nz + TZ - 1) / TZ) * a.g.B * (a.groups > 1 ? a.groups : 1));
  const size_t halo = (size_t)HALO * CIN * 2, wts = (size_t)NT * 16 * a.kp * 2;
  const size_t lim = (size_t)lds_limit();
  // weights join the halo in LDS on small grids, or where the block still fits
  // five to a CU; on large grids a larger block would cost more occupancy than
  // the cached weight reads it saves
  const bool small = (size_t)grid.x * grid.y * grid.z <= 1024;
  if (halo + wts <= lim && (small || halo + wts <= 32768)) {
    hipLaunchKernelGGL((k_conv_tiled<CIN, NT, true>), grid, dim3(256), halo + wts, st, a);
    return true;
  }
  if (halo <= lim) {
    hipLaunchKernelGGL((k_conv_tiled<CIN, NT, false>), grid, dim3(256), halo, st, a);
    return true;
  }
  return false;
}
// A grid of fewer spatial tiles than twice the CUs gets one block per
// 16-channel output tile (the input halo is then read once per tile).
template <int CIN>
static bool launch_tiled_cin(const ConvArgs& a0, hipStream_t st) {
  ConvArgs a = a0;
  const int tiles = (a.cout + 15) / 16;
  const int64_t spatial = (int64_t)((a.g.nx + TX - 1) / TX) * ((a.g.ny + TY - 1) / TY) * ((a.g.nz + TZ - 1) / TZ) * a.g.B;
  if (tiles > 1 && spatial < 512) { a.groups = tiles; return launch_tiled_nt<CIN, 1>(a, st); }
  a.groups = 1;
  switch (tiles) {
  case 1: return launch_tiled_nt<CIN, 1>(a, st);
  case 2: return launch_tiled_nt<CIN, 2>(a, st);
  case 4: return launch_tiled_nt<CIN, 4>(a, st);
  default: return false;
  }
}
// forward layers (bias column present) of every input width of the network
static bool launch_tiled(const ConvArgs& a, hipStream_t st) {
  if (!a.bias) return false;
  switch (a.ca + a.cb) {
  case 8: return launch_tiled_cin<8>(a, st);
  case 16: return launch_tiled_cin<16>(a, st);
  case 24: return launch_tiled_cin<24>(a, st);
  case 32: return launch_tiled_cin<32>(a, st);
  case 48: return launch_tiled_cin<48>(a, st);
  case 64: return launch_tiled_cin<64>(a, st);
  case 96: return launch_tiled_cin<96>(a, st);
  default: return false;
  }
}

// every input width (and padded output width, the input gradient's input) is
// a multiple of 8, checked by wv_net_create: a 32-wide K fragment stays within
// one tap and one input tensor
static void launch_conv(const ConvArgs& a, hipStream_t st) {
  if (launch_tiled(a, st)) return;
  const unsigned g = nblk(cells_of(a.g), 64);
  switch ((a.cout + 15) / 16) {
  case 1: hipLaunchKernelGGL(k_conv<1>, dim3(g), dim3(256), 0, st, a); break;
  case 2: hipLaunchKernelGGL(k_conv<2>, dim3(g), dim3(256), 0, st, a); break;
  case 3: hipLaunchKernelGGL(k_conv<3>, dim3(g), dim3(256), 0, st, a); break;
  case 4: hipLaunchKernelGGL(k_conv<4>, dim3(g), dim3(256), 0, st, a); break;
  default: hipLaunchKernelGGL(k_conv<6>, dim3(g), dim3(256), 0, st, a); break;   // up to 96 channels
  }
}

// ------------------------------------------------------- weight gradient --
#define WG_CHUNK 1024
struct WgradArgs {
  const bf16_t* dy; const bf16_t* xa; const bf16_t* xb;
  float* partial;
  Grid g, gb;
  int dys, ca, cb, kp, cout;
};
// partial[chunk][co][k] = sum over the chunk's cells of dy[c][co] A[c][k]
__global__ void __launch_bounds__(256) k_wgrad(WgradArgs a) {
  __shared__ v4f red[3][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m = lane & 15, kq = lane >> 4;
  const int64_t M = gcells(a.g);
  const int chunk = blockIdx.x, kt = blockIdx.y, ct = blockIdx.z;
  const int cin = a.ca + a.cb, kreal = 27 * cin;
  const int kidx = kt * 16 + m, co_a = ct * 16 + m;
  const bool kin = kidx < kreal, kbias = kidx == kreal, coin = co_a < a.cout;
  int tap = 0, ch = 0;
  if (kin) { tap = kidx / cin; ch = kidx - tap * cin; }
  const int dx = tap % 3 - 1, dy = (tap / 3) % 3 - 1, dz = tap / 9 - 1;
  v4f acc = {0.f, 0.f, 0.f, 0.f};
  const int64_t base = (int64_t)chunk * WG_CHUNK + wave * (WG_CHUNK / 4);
  for (int s = 0; s < WG_CHUNK / 4; s += 16) {
    v4s av, bv;
    for (int j = 0; j < 4; j++) {
      const int64_t c = base + s + 4 * kq + j;
      short ae = 0, be = 0;
      if (c < M) {
        if (coin) ae = (short)a.dy[c * a.dys + co_a];
        if (kin) {
          int x, y, z, b; decode(c, a.g, x, y, z, b);
          const int qx = x + dx, qy = y + dy, qz = z + dz;
          if (inside(a.g, qx, qy, qz))
            be = (short)(ch < a.ca ? a.xa[encode(a.g, qx, qy, qz, b) * a.ca + ch]
                                   : a.xb[encode(a.gb, qx >> 1, qy >> 1, qz >> 1, b) * a.cb + (ch - a.ca)]);
        } else if (kbias) be = (short)0x3f80;
      }
      av[j] = ae; bv[j] = be;
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
  }
  if (wave > 0) red[wave - 1][lane] = acc;
  __syncthreads();
  if (wave != 0) return;
  for (int w = 0; w < 3; w++) { v4f o = red[w][lane]; acc[0] += o[0]; acc[1] += o[1]; acc[2] += o[2]; acc[3] += o[3]; }
  for (int j = 0; j < 4; j++) {
    const int co = ct * 16 + 4 * kq + j, kk = kt * 16 + m;
    if (co < a.cout && kk < a.kp) a.partial[((int64_t)chunk * a.cout + co) * a.kp + kk] = acc[j];
  }
}
__global__ void __launch_bounds__(256) k_wgrad_sum(const float* partial, int chunks, int64_t n, float* dw) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  for (int c = 0; c < chunks; c++) s += partial[(int64_t)c * n + i];
  dw[i] = s;
}

// ---------------------------------------------------- pooling and routing --
__device__ __forceinline__ int child_count(const Grid& f, int X, int Y, int Z) {
  return (2 * X + 1 < f.nx ? 2 : 1) * (2 * Y + 1 < f.ny ? 2 : 1) * (2 * Z + 1 < f.nz ? 2 : 1);
}
// coarse[C][ch] = mean of the existing 2x2x2 children
__global__ void __launch_bounds__(256) k_pool(const bf16_t* fine, Grid f, Grid cg, int C, bf16_t* coarse) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= gcells(cg) * C) return;
  const int ch = (int)(i % C); const int64_t cc = i / C;
  int X, Y, Z, b; decode(cc, cg, X, Y, Z, b);
  float s = 0.f; int n = 0;
  for (int dz = 0; dz < 2; dz++) for (int dy = 0; dy < 2; dy++) for (int dx = 0; dx < 2; dx++) {
    const int x = 2 * X + dx, y = 2 * Y + dy, z = 2 * Z + dz;
    if (inside(f, x, y, z)) { s += bf2f(fine[encode(f, x, y, z, b) * C + ch]); n++; }
  }
  coarse[i] = f2bf(s / (float)n);
}
// dfine[c][ch] += dcoarse[parent][ch] / children(parent); dcoarse has stride ds
__global__ void __launch_bounds__(256) k_pool_bwd(const float* dcoarse, int ds, Grid f, Grid cg, int C, float* dfine) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= gcells(f) * C) return;
  const int ch = (int)(i % C); const int64_t c = i / C;
  int x, y, z, b; decode(c, f, x, y, z, b);
  const int X = x >> 1, Y = y >> 1, Z = z >> 1;
  dfine[i] += dcoarse[encode(cg, X, Y, Z, b) * ds + ch] / (float)child_count(f, X, Y, Z);
}
// skip part of a concat input gradient: dskip[c][ch] += dx[c][ch], ch < ca (dx stride cin)
__global__ void __launch_bounds__(256) k_add_skip(const float* dx, int cin, int64_t M, int ca, float* dskip) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= M * ca) return;
  const int ch = (int)(i % ca); const int64_t c = i / ca;
  dskip[i] += dx[c * cin + ch];
}
// up part: dcoarse[C][ch] += sum over fine children of dx[child][off + ch]
__global__ void __launch_bounds__(256) k_up_bwd(const float* dx, int cin, int off, Grid f, Grid cg, int cb, int factor, float* dcoarse, int dcs) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= gcells(cg) * cb) return;
  const int ch = (int)(i % cb); const int64_t cc = i / cb;
  int X, Y, Z, b; decode(cc, cg, X, Y, Z, b);
  float s = 0.f;
  for (int dz = 0; dz < factor; dz++) for (int dy = 0; dy < factor; dy++) for (int dxx = 0; dxx < factor; dxx++) {
    const int x = factor * X + dxx, y = factor * Y + dy, z = factor * Z + dz;
    if (inside(f, x, y, z)) s += dx[encode(f, x, y, z, b) * cin + off + ch];
  }
  dcoarse[cc * dcs + ch] += s;
}
// dpre = (act > 0 ? dact : 0) rounded to bf16 (stride C)
__global__ void __launch_bounds__(256) k_relu_bwd(const bf16_t* act, const float* dact, int64_t n, bf16_t* dpre) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  dpre[i] = bf2f(act[i]) > 0.f ? f2bf(dact[i]) : (bf16_t)0;
}
__global__ void __launch_bounds__(256) k_zero(float* p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = 0.f;
}

// -------------------------------------------------------------------- loss --
// per sample: num = sum_wet (q - t)^2, den = sum_wet t^2 with q = qf + up4 qc
__device__ __forceinline__ double block_sum(double v) {
  __shared__ double sh[256];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) { if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o]; __syncthreads(); }
  double r = sh[0];
  __syncthreads();
  return r;
}
__device__ __forceinline__ float q_at(const float* qf, const float* qc, const Grid& g, const Grid& g2, int64_t c) {
  int x, y, z, b; decode(c, g, x, y, z, b);
  return qf[c] + qc[encode(g2, x >> 2, y >> 2, z >> 2, b)];
}
#define LOSS_BLK 256
__global__ void __launch_bounds__(256) k_loss_partial(const float* qf, const float* qc, const float* t, const bf16_t* x,
                                                       Grid g, Grid g2, int blocks, double* part) {
  const int s = blockIdx.y;
  const int64_t per = (int64_t)g.nx * g.ny * g.nz;
  double num = 0, den = 0;
  for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < per; k += (int64_t)blocks * 256) {
    const int64_t c = s * per + k;
    if (bf2f(x[c * 8 + 3]) > 0.5f) {
      const double d = (double)q_at(qf, qc, g, g2, c) - (double)t[c];
      num += d * d; den += (double)t[c] * t[c];
    }
  }
  num = block_sum(num); den = block_sum(den);
  if (threadIdx.x == 0) { part[2 * (s * blocks + blockIdx.x)] = num; part[2 * (s * blocks + blockIdx.x) + 1] = den; }
}
// one block: per-sample den (stored in dens) and the mean loss
__global__ void __launch_bounds__(256) k_loss_final(const double* part, int blocks, int B, double* dens, float* loss) {
  double total = 0;
  for (int s = 0; s < B; s++) {
    double num = 0, den = 0;
    for (int k = threadIdx.x; k < blocks; k += 256) { num += part[2 * (s * blocks + k)]; den += part[2 * (s * blocks + k) + 1]; }
    num = block_sum(num); den = block_sum(den);
    den += 1e-30;
    if (threadIdx.x == 0) dens[s] = den;
    total += num / den;
  }
  if (threadIdx.x == 0) *loss = (float)(total / B);
}
// dq = 2 (q - t) wet / den / B -> dqf (fp32) and the fine head's dpre (bf16, stride 8)
__global__ void __launch_bounds__(256) k_loss_grad(const float* qf, const float* qc, const float* t, const bf16_t* x,
                                                    Grid g, Grid g2, const double* dens, float* dq, bf16_t* dpre) {
  const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= gcells(g)) return;
  const int64_t per = (int64_t)g.nx * g.ny * g.nz;
  const int s = (int)(c / per);
  float v = 0.f;
  if (bf2f(x[c * 8 + 3]) > 0.5f) v = (float)(2.0 * ((double)q_at(qf, qc, g, g2, c) - (double)t[c]) / dens[s] / g.B);
  dq[c] = v;
  dpre[c * 8] = f2bf(v);
  for (int k = 1; k < 8; k++) dpre[c * 8 + k] = 0;
}
// coarse head: dpre[C*8] = sum of dq over the fine cells that read C (factor 4)
__global__ void __launch_bounds__(256) k_head_coarse_bwd(const float* dq, Grid g, Grid g2, bf16_t* dpre) {
  const int64_t cc = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (cc >= gcells(g2)) return;
  int X, Y, Z, b; decode(cc, g2, X, Y, Z, b);
  float s = 0.f;
  for (int dz = 0; dz < 4; dz++) for (int dy = 0; dy < 4; dy++) for (int dx = 0; dx < 4; dx++) {
    const int x = 4 * X + dx, y = 4 * Y + dy, z = 4 * Z + dz;
    if (inside(g, x, y, z)) s += dq[encode(g, x, y, z, b)];
  }
  dpre[cc * 8] = f2bf(s);
  for (int k = 1; k < 8; k++) dpre[cc * 8 + k] = 0;
}

// ------------------------------------------------------------ parameters --
__global__ void __launch_bounds__(256) k_adam(float* w, float* mo, float* ve, const float* g, int64_t n,
                                               float lr, float b1, float b2, float eps, float c1, float c2) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float m = b1 * mo[i] + (1.f - b1) * gi;
  const float v = b2 * ve[i] + (1.f - b2) * gi * gi;
  mo[i] = m; ve[i] = v;
  w[i] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}
__global__ void __launch_bounds__(256) k_to_bf16(const float* w, int64_t n, bf16_t* out) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i] = f2bf(w[i]);
}
// Wt[ci][t coutp + co] = W[co][(26 - t) cin + ci] (co < cout), zero elsewhere
__global__ void __launch_bounds__(256) k_transpose(const float* w, int cout, int cin, int kp, int coutp, int kpt, bf16_t* wt) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (int64_t)cin * kpt) return;
  const int ci = (int)(i / kpt), k = (int)(i % kpt);
  float v = 0.f;
  if (k < 27 * coutp) {
    const int t = k / coutp, co = k % coutp;
    if (co < cout) v = w[(int64_t)co * kp + (26 - t) * cin + ci];
  }
  wt[i] = f2bf(v);
}

// ================================================================== host ==
namespace {
enum { SRC_PLAIN = 0, SRC_POOL = 1, SRC_CAT = 2 };
struct LayerDef { const char* name; int level; int kind; int src; int src2; int ca; int cb; int cout; int relu; };
// src / src2 index into this table; -1 = the network input
const LayerDef DEFS[] = {
    {"e0a", 0, SRC_PLAIN, -1, -1, 8, 0, 8, 1},
    {"e0b", 0, SRC_PLAIN, 0, -1, 8, 0, 8, 1},
    {"e1a", 1, SRC_POOL, 1, -1, 8, 0, 16, 1},
    {"e1b", 1, SRC_PLAIN, 2, -1, 16, 0, 16, 1},
    {"e2a", 2, SRC_POOL, 3, -1, 16, 0, 32, 1},
    {"e2b", 2, SRC_PLAIN, 4, -1, 32, 0, 32, 1},
    {"b3a", 3, SRC_POOL, 5, -1, 32, 0, 64, 1},
    {"b3b", 3, SRC_PLAIN, 6, -1, 64, 0, 64, 1},
    {"d2a", 2, SRC_CAT, 5, 7, 32, 64, 32, 1},
    {"d2b", 2, SRC_PLAIN, 8, -1, 32, 0, 32, 1},
    {"hc", 2, SRC_PLAIN, 9, -1, 32, 0, 1, 0},
    {"d1a", 1, SRC_CAT, 3, 9, 16, 32, 16, 1},
    {"d1b", 1, SRC_PLAIN, 11, -1, 16, 0, 16, 1},
    {"d0a", 0, SRC_CAT, 1, 12, 8, 16, 8, 1},
    {"d0b", 0, SRC_PLAIN, 13, -1, 8, 0, 8, 1},
    {"hf", 0, SRC_PLAIN, 14, -1, 8, 0, 1, 0},
};
const int NL = sizeof(DEFS) / sizeof(DEFS[0]);
const int HC = 10, HF = 15;
inline int kpad(int cin) { return (27 * cin + 1 + 15) / 16 * 16; }
inline int coutp_of(int cout) { return cout < 8 ? 8 : cout; }
}

struct wv_net {
  Grid lv[4];
  int batch = 0;
  hipStream_t stream = nullptr;
  std::string err;
  std::vector<void*> allocs;
  int64_t woff[NL], wtoff[NL], nparams = 0, nwt = 0;
  float *w = nullptr, *g = nullptr, *mo = nullptr, *ve = nullptr;   // fp32 master, grad, Adam moments
  bf16_t *wb = nullptr, *wt = nullptr;                              // bf16 weights, transposed
  bf16_t* x = nullptr;                                              // input [cells][8]
  bf16_t* act[NL] = {};                                             // relu layers [cells][cout]
  bf16_t* pooled[NL] = {};                                          // pooled inputs of SRC_POOL layers
  float *qf = nullptr, *qc = nullptr;                               // head outputs
  float* dact[NL] = {};                                             // activation gradients
  bf16_t* dpre = nullptr;                                           // scratch: pre-activation gradient
  float* dx = nullptr;                                              // scratch: input gradient
  float* partial = nullptr; int64_t partial_n = 0;
  float* t = nullptr; float* dq = nullptr;                          // target batch, d loss / d q
  double *lpart = nullptr, *dens = nullptr; int lblocks = 0;
  float* dloss = nullptr;                                           // per-step losses
  int loss_cap = 0;
  bf16_t* data_x = nullptr; float* data_t = nullptr; int data_count = 0;
  int adam_t = 0;
};

static int net_fail(wv_net* n, const char* what, hipError_t e) {
  n->err = std::string(what) + ": " + hipGetErrorString(e);
  return WV_EHIP;
}
#define NCHECK(n, call) do { hipError_t e_ = (call); if (e_ != hipSuccess) return net_fail(n, #call, e_); } while (0)

template <class T> static int nalloc(wv_net* n, T** p, int64_t count) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, (size_t)count * sizeof(T) + 256);
  if (e != hipSuccess) return net_fail(n, "hipMalloc", e);
  n->allocs.push_back(q);
  (void)hipMemset(q, 0, (size_t)count * sizeof(T) + 256);
  *p = (T*)q;
  return WV_OK;
}

static int64_t level_cells(const wv_net* n, int l) { return cells_of(n->lv[l]); }

extern "C" int64_t wv_net_param_count(void) {
  int64_t s = 0;
  for (int l = 0; l < NL; l++) s += (int64_t)DEFS[l].cout * kpad(DEFS[l].ca + DEFS[l].cb);
  return s;
}

extern "C" void wv_net_destroy(wv_net* n) {
  if (!n) return;
  if (n->stream) (void)hipStreamSynchronize(n->stream);
  for (void* p : n->allocs) (void)hipFree(p);
  if (n->stream) (void)hipStreamDestroy(n->stream);
  delete n;
}

extern "C" int wv_net_create(int nx, int ny, int nz, int batch, wv_net** out) {
  if (!out || nx < 1 || ny < 1 || nz < 1 || batch < 1) return WV_EINVAL;
  *out = nullptr;
  for (int l = 0; l < NL; l++)
    if (DEFS[l].ca % 8 != 0 || DEFS[l].cb % 8 != 0 || coutp_of(DEFS[l].cout) % 8 != 0) return WV_EINVAL;
  wv_net* n = new wv_net();
  n->batch = batch;
  n->lv[0] = Grid{nx, ny, nz, batch};
  for (int l = 1; l < 4; l++) n->lv[l] = Grid{(n->lv[l - 1].nx + 1) / 2, (n->lv[l - 1].ny + 1) / 2, (n->lv[l - 1].nz + 1) / 2, batch};
  int rc = WV_OK;
  if (hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess) { delete n; return WV_EHIP; }
  for (int l = 0; l < NL; l++) {
    const int cin = DEFS[l].ca + DEFS[l].cb;
    n->woff[l] = n->nparams; n->nparams += (int64_t)DEFS[l].cout * kpad(cin);
    n->wtoff[l] = n->nwt; n->nwt += (int64_t)cin * kpad(coutp_of(DEFS[l].cout));
  }
#define NA(p, c) if ((rc = nalloc(n, &(p), (c))) != WV_OK) { wv_net_destroy(n); return rc; }
  NA(n->w, n->nparams); NA(n->g, n->nparams); NA(n->mo, n->nparams); NA(n->ve, n->nparams);
  NA(n->wb, n->nparams); NA(n->wt, n->nwt);
  const int64_t M0 = level_cells(n, 0);
  NA(n->x, M0 * 8);
  int64_t max_pre = 0, max_dx = 0, max_part = 0;
  for (int l = 0; l < NL; l++) {
    const LayerDef& d = DEFS[l];
    const int64_t M = level_cells(n, d.level);
    const int cin = d.ca + d.cb, cp = coutp_of(d.cout);
    if (d.relu) NA(n->act[l], M * d.cout);
    if (d.kind == SRC_POOL) NA(n->pooled[l], M * d.ca);
    NA(n->dact[l], M * (d.relu ? d.cout : 1));
    max_pre = std::max(max_pre, M * cp);
    max_dx = std::max(max_dx, M * cin);
    max_part = std::max(max_part, (int64_t)((M + WG_CHUNK - 1) / WG_CHUNK) * d.cout * kpad(cin));
  }
  NA(n->qf, M0); NA(n->qc, level_cells(n, 2));
  NA(n->dpre, max_pre); NA(n->dx, max_dx); NA(n->partial, max_part);
  n->partial_n = max_part;
  NA(n->t, M0); NA(n->dq, M0);
  n->lblocks = 64;
  NA(n->lpart, 2 * (int64_t)batch * n->lblocks); NA(n->dens, batch);
#undef NA
  *out = n;
  return WV_OK;
}

static void refresh_weights(wv_net* n) {
  hipStream_t st = n->stream;
  hipLaunchKernelGGL(k_to_bf16, dim3(nblk(n->nparams, 256)), dim3(256), 0, st, n->w, n->nparams, n->wb);
  for (int l = 0; l < NL; l++) {
    const LayerDef& d = DEFS[l];
    const int cin = d.ca + d.cb, cp = coutp_of(d.cout), kpt = kpad(cp);
    hipLaunchKernelGGL(k_transpose, dim3(nblk((int64_t)cin * kpt, 256)), dim3(256), 0, st,
                       n->w + n->woff[l], d.cout, cin, kpad(cin), cp, kpt, n->wt + n->wtoff[l]);
  }
}

extern "C" int wv_net_set_params(wv_net* n, const float* p, int64_t count) {
  if (!n || !p || count != n->nparams) return WV_EINVAL;
  NCHECK(n, hipMemcpyAsync(n->w, p, count * 4, hipMemcpyHostToDevice, n->stream));
  NCHECK(n, hipMemsetAsync(n->mo, 0, count * 4, n->stream));
  NCHECK(n, hipMemsetAsync(n->ve, 0, count * 4, n->stream));
  n->adam_t = 0;
  refresh_weights(n);
  NCHECK(n, hipStreamSynchronize(n->stream));
  return WV_OK;
}
extern "C" int wv_net_get_params(wv_net* n, float* p, int64_t count) {
  if (!n || !p || count != n->nparams) return WV_EINVAL;
  NCHECK(n, hipMemcpyAsync(p, n->w, count * 4, hipMemcpyDeviceToHost, n->stream));
  NCHECK(n, hipStreamSynchronize(n->stream));
  return WV_OK;
}

// input tensor of layer l (skip part) and its grid
static const bf16_t* layer_input(wv_net* n, int l) {
  const LayerDef& d = DEFS[l];
  if (d.kind == SRC_POOL) return n->pooled[l];
  return d.src < 0 ? n->x : n->act[d.src];
}

// the forward convolution of layer l (its inputs computed)
static void forward_conv(wv_net* n, int l, hipStream_t st);
static void forward(wv_net* n, hipStream_t st) {
  for (int l = 0; l < NL; l++) {
    const LayerDef& d = DEFS[l];
    if (d.kind == SRC_POOL) {
      const Grid f = n->lv[d.level - 1], c = n->lv[d.level];
      hipLaunchKernelGGL(k_pool, dim3(nblk(cells_of(c) * d.ca, 256)), dim3(256), 0, st, n->act[d.src], f, c, d.ca, n->pooled[l]);
    }
    forward_conv(n, l, st);
  }
}
static void forward_conv(wv_net* n, int l, hipStream_t st) {
  const LayerDef& d = DEFS[l];
  ConvArgs a{};
  a.xa = layer_input(n, l);
  a.xb = d.kind == SRC_CAT ? n->act[d.src2] : nullptr;
  a.w = n->wb + n->woff[l];
  a.g = n->lv[d.level];
  a.gb = d.kind == SRC_CAT ? n->lv[d.level + 1] : a.g;
  a.ca = d.ca; a.cb = d.cb; a.kp = kpad(d.ca + d.cb); a.cout = d.cout; a.bias = 1; a.relu = d.relu;
  if (d.relu) { a.yb = n->act[l]; a.ystride = d.cout; }
  else { a.yf = l == HF ? n->qf : n->qc; a.ystride = 1; }
  launch_conv(a, st);
}

// loss of the current batch (x, t) into *dloss_slot, gradients into n->g
static void loss_backward(wv_net* n, float* dloss_slot, bool backward, hipStream_t st) {
  const Grid g0 = n->lv[0], g2 = n->lv[2];
  hipLaunchKernelGGL(k_loss_partial, dim3(n->lblocks, n->batch), dim3(256), 0, st, n->qf, n->qc, n->t, n->x, g0, g2, n->lblocks, n->lpart);
  hipLaunchKernelGGL(k_loss_final, dim3(1), dim3(256), 0, st, n->lpart, n->lblocks, n->batch, n->dens, dloss_slot);
  if (!backward) return;
  for (int l = 0; l < NL; l++) {
    const int64_t M = level_cells(n, DEFS[l].level) * (DEFS[l].relu ? DEFS[l].cout : 1);
    hipLaunchKernelGGL(k_zero, dim3(nblk(M, 256)), dim3(256), 0, st, n->dact[l], M);
  }
  for (int l = NL - 1; l >= 0; l--) {
    const LayerDef& d = DEFS[l];
    const Grid g = n->lv[d.level];
    const int64_t M = cells_of(g);
    const int cin = d.ca + d.cb, cp = coutp_of(d.cout);
    // pre-activation gradient (bf16) into dpre, stride cp
    if (l == HF) {
      hipLaunchKernelGGL(k_loss_grad, dim3(nblk(M, 256)), dim3(256), 0, st, n->qf, n->qc, n->t, n->x, g0, g2, n->dens, n->dq, n->dpre);
    } else if (l == HC) {
      hipLaunchKernelGGL(k_head_coarse_bwd, dim3(nblk(M, 256)), dim3(256), 0, st, n->dq, g0, g2, n->dpre);
    } else {
      hipLaunchKernelGGL(k_relu_bwd, dim3(nblk(M * d.cout, 256)), dim3(256), 0, st, n->act[l], n->dact[l], M * d.cout, n->dpre);
    }
    // weight gradient
    WgradArgs wa{};
    wa.dy = n->dpre; wa.dys = cp;
    wa.xa = layer_input(n, l);
    wa.xb = d.kind == SRC_CAT ? n->act[d.src2] : nullptr;
    wa.g = g; wa.gb = d.kind == SRC_CAT ? n->lv[d.level + 1] : g;
    wa.ca = d.ca; wa.cb = d.cb; wa.kp = kpad(cin); wa.cout = d.cout; wa.partial = n->partial;
    const int chunks = (int)((M + WG_CHUNK - 1) / WG_CHUNK);
    hipLaunchKernelGGL(k_wgrad, dim3(chunks, wa.kp / 16, (d.cout + 15) / 16), dim3(256), 0, st, wa);
    const int64_t nw = (int64_t)d.cout * wa.kp;
    hipLaunchKernelGGL(k_wgrad_sum, dim3(nblk(nw, 256)), dim3(256), 0, st, n->partial, chunks, nw, n->g + n->woff[l]);
    if (d.src < 0) continue;                  // no gradient into the network input
    // input gradient: conv of dpre with the mirrored, transposed weights
    ConvArgs a{};
    a.xa = n->dpre; a.xb = nullptr; a.w = n->wt + n->wtoff[l];
    a.g = g; a.gb = g; a.ca = cp; a.cb = 0; a.kp = kpad(cp); a.cout = cin; a.bias = 0; a.relu = 0;
    a.yf = n->dx; a.ystride = cin;
    launch_conv(a, st);
    if (d.kind == SRC_PLAIN) {
      hipLaunchKernelGGL(k_add_skip, dim3(nblk(M * cin, 256)), dim3(256), 0, st, n->dx, cin, M, cin, n->dact[d.src]);
    } else if (d.kind == SRC_POOL) {
      const Grid f = n->lv[d.level - 1];
      hipLaunchKernelGGL(k_pool_bwd, dim3(nblk(cells_of(f) * cin, 256)), dim3(256), 0, st, n->dx, cin, f, g, cin, n->dact[d.src]);
    } else {
      const Grid c = n->lv[d.level + 1];
      hipLaunchKernelGGL(k_add_skip, dim3(nblk(M * d.ca, 256)), dim3(256), 0, st, n->dx, cin, M, d.ca, n->dact[d.src]);
      hipLaunchKernelGGL(k_up_bwd, dim3(nblk(cells_of(c) * d.cb, 256)), dim3(256), 0, st, n->dx, cin, d.ca, g, c, d.cb, 2,
                         n->dact[d.src2], d.cb);
    }
  }
}

extern "C" int wv_net_layer_count(void) { return NL; }
extern "C" int wv_net_layer_info(int l, int* level, int* ca, int* cb, int* cout) {
  if (l < 0 || l >= NL || !level || !ca || !cb || !cout) return WV_EINVAL;
  *level = DEFS[l].level; *ca = DEFS[l].ca; *cb = DEFS[l].cb; *cout = DEFS[l].cout;
  return WV_OK;
}
extern "C" int wv_net_time_layers(wv_net* n, int reps, float* us, int count) {
  if (!n || !us || count < NL || reps < 1) return WV_EINVAL;
  forward(n, n->stream);
  hipEvent_t e0, e1;
  NCHECK(n, hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return net_fail(n, "hipEventCreate", hipErrorUnknown); }
  int rc = WV_OK;
  for (int l = 0; l < NL && rc == WV_OK; l++) {
    forward_conv(n, l, n->stream);                          // warm
    hipError_t e = hipEventRecord(e0, n->stream);
    for (int r = 0; r < reps && e == hipSuccess; r++) forward_conv(n, l, n->stream);
    if (e == hipSuccess) e = hipEventRecord(e1, n->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e != hipSuccess) rc = net_fail(n, "layer timing", e);
    else us[l] = ms * 1000.0f / (float)reps;
  }
  (void)hipEventDestroy(e0); (void)hipEventDestroy(e1);
  return rc;
}

extern "C" int wv_net_forward(wv_net* n, const uint16_t* x, float* qf, float* qc) {
  if (!n || !x) return WV_EINVAL;
  const int64_t M0 = level_cells(n, 0), M2 = level_cells(n, 2);
  NCHECK(n, hipMemcpyAsync(n->x, x, M0 * 8 * 2, hipMemcpyHostToDevice, n->stream));
  forward(n, n->stream);
  NCHECK(n, hipGetLastError());
  if (qf) NCHECK(n, hipMemcpyAsync(qf, n->qf, M0 * 4, hipMemcpyDeviceToHost, n->stream));
  if (qc) NCHECK(n, hipMemcpyAsync(qc, n->qc, M2 * 4, hipMemcpyDeviceToHost, n->stream));
  NCHECK(n, hipStreamSynchronize(n->stream));
  return WV_OK;
}

static int ensure_loss_slots(wv_net* n, int count) {
  if (count <= n->loss_cap) return WV_OK;
  float* p = nullptr;
  int rc = nalloc(n, &p, count);
  if (rc) return rc;
  n->dloss = p; n->loss_cap = count;
  return WV_OK;
}

extern "C" int wv_net_loss_grad(wv_net* n, const uint16_t* x, const float* t, float* loss, float* grad) {
  if (!n || !x || !t) return WV_EINVAL;
  int rc = ensure_loss_slots(n, 1); if (rc) return rc;
  const int64_t M0 = level_cells(n, 0);
  NCHECK(n, hipMemcpyAsync(n->x, x, M0 * 8 * 2, hipMemcpyHostToDevice, n->stream));
  NCHECK(n, hipMemcpyAsync(n->t, t, M0 * 4, hipMemcpyHostToDevice, n->stream));
  forward(n, n->stream);
  loss_backward(n, n->dloss, grad != nullptr, n->stream);
  NCHECK(n, hipGetLastError());
  if (loss) NCHECK(n, hipMemcpyAsync(loss, n->dloss, 4, hipMemcpyDeviceToHost, n->stream));
  if (grad) NCHECK(n, hipMemcpyAsync(grad, n->g, n->nparams * 4, hipMemcpyDeviceToHost, n->stream));
  NCHECK(n, hipStreamSynchronize(n->stream));
  return WV_OK;
}

extern "C" int wv_net_set_dataset(wv_net* n, const uint16_t* x, const float* t, int count) {
  if (!n || !x || !t || count < 1) return WV_EINVAL;
  const int64_t per = cells_of(Grid{n->lv[0].nx, n->lv[0].ny, n->lv[0].nz, 1});
  if (count > n->data_count) {
    int rc;
    if ((rc = nalloc(n, &n->data_x, per * 8 * count)) != WV_OK) return rc;
    if ((rc = nalloc(n, &n->data_t, per * count)) != WV_OK) return rc;
  }
  n->data_count = count;
  NCHECK(n, hipMemcpyAsync(n->data_x, x, per * 8 * 2 * count, hipMemcpyHostToDevice, n->stream));
  NCHECK(n, hipMemcpyAsync(n->data_t, t, per * 4 * count, hipMemcpyHostToDevice, n->stream));
  NCHECK(n, hipStreamSynchronize(n->stream));
  return WV_OK;
}

static inline uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static int load_batch(wv_net* n, const int* idx) {
  const int64_t per = cells_of(Grid{n->lv[0].nx, n->lv[0].ny, n->lv[0].nz, 1});
  for (int b = 0; b < n->batch; b++) {
    NCHECK(n, hipMemcpyAsync(n->x + per * 8 * b, n->data_x + per * 8 * idx[b], per * 8 * 2, hipMemcpyDeviceToDevice, n->stream));
    NCHECK(n, hipMemcpyAsync(n->t + per * b, n->data_t + per * idx[b], per * 4, hipMemcpyDeviceToDevice, n->stream));
  }
  return WV_OK;
}

extern "C" int wv_net_train(wv_net* n, int steps, int train_count, float lr, uint64_t seed, float* losses) {
  if (!n || steps < 1 || train_count < 1 || train_count > n->data_count || !(lr > 0)) return WV_EINVAL;
  int rc = ensure_loss_slots(n, steps); if (rc) return rc;
  uint64_t s = seed;
  std::vector<int> idx(n->batch);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int k = 0; k < steps; k++) {
    for (int b = 0; b < n->batch; b++) idx[b] = (int)(splitmix(s) % (uint64_t)train_count);
    if ((rc = load_batch(n, idx.data())) != WV_OK) return rc;
    forward(n, n->stream);
    loss_backward(n, n->dloss + k, true, n->stream);
    n->adam_t++;
    const float c1 = 1.f - powf(b1, (float)n->adam_t), c2 = 1.f - powf(b2, (float)n->adam_t);
    hipLaunchKernelGGL(k_adam, dim3(nblk(n->nparams, 256)), dim3(256), 0, n->stream, n->w, n->mo, n->ve, n->g, n->nparams,
                       lr, b1, b2, eps, c1, c2);
    refresh_weights(n);
    NCHECK(n, hipGetLastError());
    if ((k & 63) == 63) NCHECK(n, hipStreamSynchronize(n->stream));   // bound the queued work
  }
  if (losses) NCHECK(n, hipMemcpyAsync(losses, n->dloss, steps * 4, hipMemcpyDeviceToHost, n->stream));
  NCHECK(n, hipStreamSynchronize(n->stream));
  return WV_OK;
}

extern "C" int wv_net_eval(wv_net* n, int first, int count, float* loss) {
  if (!n || !loss || first < 0 || count < 1 || first + count > n->data_count) return WV_EINVAL;
  const int batches = (count + n->batch - 1) / n->batch;
  int rc = ensure_loss_slots(n, batches); if (rc) return rc;
  std::vector<int> idx(n->batch);
  for (int k = 0; k < batches; k++) {
    for (int b = 0; b < n->batch; b++) idx[b] = first + std::min(k * n->batch + b, count - 1);
    if ((rc = load_batch(n, idx.data())) != WV_OK) return rc;
    forward(n, n->stream);
    loss_backward(n, n->dloss + k, false, n->stream);
  }
  NCHECK(n, hipGetLastError());
  std::vector<float> l(batches);
  NCHECK(n, hipMemcpyAsync(l.data(), n->dloss, batches * 4, hipMemcpyDeviceToHost, n->stream));
  NCHECK(n, hipStreamSynchronize(n->stream));
  double s = 0;
  for (float v : l) s += v;
  *loss = (float)(s / batches);
  return WV_OK;
}

extern "C" const char* wv_net_last_error(wv_net* n) { return n ? n->err.c_str() : "null net"; }

// ---------------------------------------------------- solver-side access --
uint16_t* wv_net_input_dev(wv_net* n) { return n->x; }
const float* wv_net_qf_dev(wv_net* n) { return n->qf; }
const float* wv_net_qc_dev(wv_net* n) { return n->qc; }
void wv_net_grid(const wv_net* n, int* dims) { dims[0] = n->lv[0].nx; dims[1] = n->lv[0].ny; dims[2] = n->lv[0].nz; dims[3] = n->batch; }
void wv_net_forward_dev(wv_net* n, hipStream_t st) { forward(n, st); }
