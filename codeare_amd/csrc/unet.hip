// unet.hip -- the learned CG warm start: a 2D U-Net (3x3 convolutions, ReLU,
// 2x2 max-pool down, nearest 2x up + skip concatenation) trained and run with
// hand-written BF16 MFMA kernels (v_mfma_f32_32x32x16_bf16, FP32 accumulate).
//
// Tensors are NHWC: activation[((b*H + y)*W + x)*C + c], BF16 storage.  The
// network input and output are complex64 images, i.e. NHWC with C = 2
// (re, im).  Conv weights are kept as FP32 masters [tap][co][ci] (tap = 3*ky+kx)
// with a BF16 copy in that layout for the forward pass and a flipped, transposed
// BF16 copy [8-tap][ci][co] for the input-gradient pass, which is then the same
// convolution kernel.  Every reduction (weight and bias gradients, loss) runs in
// a fixed order, so a training step is deterministic.
//
// There is no counterpart in codeare: the network supplies x0 for the CG solve
// (NCSENSE::Adjoint -> CGLS::Solve, src/optimisation/CGLS.hpp:30-59, starts
// from zero).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/cgs_unet.h"
#include "../../include/cgsense.h"

namespace cgs {
void set_error(const std::string& m);
}

namespace {

using bf16 = uint16_t;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ float bf2f(bf16 h) { return __uint_as_float((uint32_t)h << 16); }
__host__ __device__ inline bf16 f2bf(float f) {
  uint32_t u;
#ifdef __HIP_DEVICE_COMPILE__
  u = __float_as_uint(f);
#else
  std::memcpy(&u, &f, 4);
#endif
  if ((u & 0x7f800000u) == 0x7f800000u) return (bf16)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));  // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (bf16)(u >> 16);
}

__device__ __forceinline__ bf16x8 as_bf8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }

// ----------------------------------------------------------------------------
// 3x3 convolution, stride 1, zero padding 1, as an implicit GEMM on MFMA.
// Input channels [0, Ca) come from A (H x W), channels [Ca, Ca+Cb) from B at
// half resolution read through a nearest 2x upsampling (B may be absent).
// Output channels [0, split) go to out0 (channel stride `split`), optionally
// multiplied by (mask0 > 0) with mask0 laid out like out0; channels
// [split, Cout) go to out1 (stride Cout - split).  bias/relu: forward epilogue.
struct ConvArgs {
  const bf16* A;
  const bf16* B;
  const bf16* Wt;     // [9][Cout][Cin]
  const float* bias;  // [Cout] or null
  bf16* out0;
  bf16* out1;
  const bf16* mask0;
  int Ca, Cb, Cout, split;
  int H, W, batch;
  int relu;
};

// Block: 4 waves; tile TR x TW pixels x BC output channels; each wave owns
// 64 pixels (two 32-pixel MFMA rows) x 32 channels.  K is walked in chunks of
// 32 input channels over all 9 taps, staged through LDS.
template <int TR, int TW, int BC>
__global__ __launch_bounds__(256) void k_conv3x3(ConvArgs a) {
  constexpr int KC = 32;
  constexpr int HS = KC + 8;            // halves per halo position (16 B pad)
  constexpr int WS = KC + 8;            // halves per weight row
  constexpr int HP = (TR + 2) * (TW + 2);
  constexpr int NWP = TR * TW / 64;     // pixel groups
  static_assert(NWP * (BC / 32) == 4, "tile must give 4 waves");
  __shared__ __attribute__((aligned(16))) bf16 hal[HP * HS];
  __shared__ __attribute__((aligned(16))) bf16 wl[9 * BC * WS];

  const int Cin = a.Ca + a.Cb;
  const int tiles_x = a.W / TW, tiles_y = a.H / TR;
  int t = blockIdx.x;
  const int tx0 = (t % tiles_x) * TW;
  t /= tiles_x;
  const int ty0 = (t % tiles_y) * TR;
  const int b = t / tiles_y;
  const int co0 = blockIdx.y * BC;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wp = w % NWP, wc = w / NWP;
  const int row = lane & 31, kh = lane >> 5;
  const int Hh = a.H >> 1, Wh = a.W >> 1;

  // pixel positions of this lane's two A rows
  int hpos[2];
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const int p = wp * 64 + s * 32 + row;
    hpos[s] = (p / TW) * (TW + 2) + (p % TW);
  }

  f32x16 acc[2];
#pragma unroll
  for (int s = 0; s < 2; ++s)
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[s][r] = 0.f;

  for (int c0 = 0; c0 < Cin; c0 += KC) {
    const int kc = min(KC, Cin - c0);   // 16 or 32
    const int g = kc >> 3;
    __syncthreads();
    // ---- stage the input halo: HP positions x g groups of 8 channels
    {
      const int total = HP * g;
      for (int base = 0; base < total; base += 256 * 4) {
        uint4 v[4];
        int dst[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int idx = base + u * 256 + tid;
          dst[u] = -1;
          v[u] = make_uint4(0, 0, 0, 0);
          if (idx < total) {
            const int pos = idx / g, grp = idx - pos * g;
            const int hy = pos / (TW + 2), hx = pos - hy * (TW + 2);
            const int y = ty0 + hy - 1, x = tx0 + hx - 1;
            const int ci = c0 + grp * 8;
            dst[u] = pos * HS + grp * 8;
            if (y >= 0 && y < a.H && x >= 0 && x < a.W) {
              const bf16* src;
              if (ci < a.Ca)
                src = a.A + ((size_t)(b * a.H + y) * a.W + x) * a.Ca + ci;
              else
                src = a.B + ((size_t)(b * Hh + (y >> 1)) * Wh + (x >> 1)) * a.Cb + (ci - a.Ca);
              v[u] = *reinterpret_cast<const uint4*>(src);
            }
          }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (dst[u] >= 0) *reinterpret_cast<uint4*>(&hal[dst[u]]) = v[u];
      }
    }
    // ---- stage the weights: 9 taps x BC channels x g groups
    {
      const int total = 9 * BC * g;
      for (int base = 0; base < total; base += 256 * 4) {
        uint4 v[4];
        int dst[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int idx = base + u * 256 + tid;
          dst[u] = -1;
          if (idx < total) {
            const int grp = idx % g, rc = idx / g;   // rc = tap*BC + co
            const int tap = rc / BC, co = rc - tap * BC;
            dst[u] = rc * WS + grp * 8;
            v[u] = *reinterpret_cast<const uint4*>(a.Wt + ((size_t)tap * a.Cout + co0 + co) * Cin + c0 + grp * 8);
          }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (dst[u] >= 0) *reinterpret_cast<uint4*>(&wl[dst[u]]) = v[u];
      }
    }
    __syncthreads();
    // ---- MFMA over 9 taps x kc/16 steps
#pragma unroll 1
    for (int tap = 0; tap < 9; ++tap) {
      const int dy = tap / 3, dx = tap - dy * 3;
      const int shift = dy * (TW + 2) + dx;
      for (int ks = 0; ks < (kc >> 4); ++ks) {
        const int koff = ks * 16 + kh * 8;
        const bf16x8 bv = as_bf8(*reinterpret_cast<const uint4*>(&wl[(tap * BC + wc * 32 + row) * WS + koff]));
#pragma unroll
        for (int s = 0; s < 2; ++s) {
          const bf16x8 av = as_bf8(*reinterpret_cast<const uint4*>(&hal[(hpos[s] + shift) * HS + koff]));
          acc[s] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc[s], 0, 0, 0);
        }
      }
    }
  }

  // ---- epilogue: D[row=pixel][col=channel]; lane holds col lane&31 and
  // rows 8*(r/4) + 4*(lane/32) + (r%4)
  const int co = co0 + wc * 32 + (lane & 31);
  const float bv = a.bias ? a.bias[co] : 0.f;
#pragma unroll
  for (int s = 0; s < 2; ++s)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int p = wp * 64 + s * 32 + 8 * (r >> 2) + 4 * kh + (r & 3);
      const int y = ty0 + p / TW, x = tx0 + p % TW;
      const size_t pix = (size_t)(b * a.H + y) * a.W + x;
      float v = acc[s][r] + bv;
      if (a.relu) v = fmaxf(v, 0.f);
      if (co < a.split) {
        const size_t o = pix * a.split + co;
        if (a.mask0 && !(bf2f(a.mask0[o]) > 0.f)) v = 0.f;
        a.out0[o] = f2bf(v);
      } else {
        a.out1[pix * (a.Cout - a.split) + (co - a.split)] = f2bf(v);
      }
    }
}

// ----------------------------------------------------------------------------
// Weight gradient of the 3x3 convolution: dW[tap][co][ci] = sum over pixels of
// dY[p][co] * X[p + off(tap)][ci], X = the forward input (A | up(B)).  The
// pixels are walked in 8x8 tiles; blockIdx.z takes a contiguous range of tiles
// and writes its partial sums (part[z][tap][co][ci]); k_reduce_splits adds the
// ranges in a fixed order.  Block: COT x CIT channels, one wave per 32 x 32
// sub-tile holding all 9 taps.
struct WgradArgs {
  const bf16* dY;  // [B][H][W][Cout], gradient at the pre-activation
  const bf16* A;
  const bf16* B;
  float* part;
  int Ca, Cb, Cout, H, W, batch;
  int tiles_per_split, ntiles;
};

template <int COT, int CIT>
__global__ __launch_bounds__(COT * CIT / 16) void k_conv3x3_wgrad(WgradArgs a) {
  constexpr int NT = COT * CIT / 16;   // threads (64 per 32x32 sub-tile)
  constexpr int DS = 64 + 8;           // halves per dY^T row (co)
  constexpr int XS = 80 + 8;           // halves per (dx, ci) row: 10 halo rows x 8 px
  __shared__ __attribute__((aligned(16))) bf16 dyT[COT * DS];
  __shared__ __attribute__((aligned(16))) bf16 xs[3 * CIT * XS];

  const int Cin = a.Ca + a.Cb;
  const int co0 = blockIdx.x * COT, ci0 = blockIdx.y * CIT;
  const int t_begin = blockIdx.z * a.tiles_per_split;
  const int t_end = min(a.ntiles, t_begin + a.tiles_per_split);
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wco = w % (COT / 32), wci = w / (COT / 32);
  const int row = lane & 31, kh = lane >> 5;
  const int tiles_x = a.W >> 3, tiles_y = a.H >> 3;
  const int Hh = a.H >> 1, Wh = a.W >> 1;

  f32x16 acc[9];
#pragma unroll
  for (int q = 0; q < 9; ++q)
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[q][r] = 0.f;

  for (int t = t_begin; t < t_end; ++t) {
    int tt = t;
    const int x0 = (tt % tiles_x) * 8;
    tt /= tiles_x;
    const int y0 = (tt % tiles_y) * 8;
    const int b = tt / tiles_y;
    __syncthreads();
    // dY tile (64 px x COT) -> dyT[co][px]
    for (int idx = tid; idx < 64 * (COT / 8); idx += NT) {
      const int p = idx / (COT / 8), grp = idx - p * (COT / 8);
      const int y = y0 + (p >> 3), x = x0 + (p & 7);
      const uint4 v = *reinterpret_cast<const uint4*>(a.dY + ((size_t)(b * a.H + y) * a.W + x) * a.Cout + co0 + grp * 8);
      const bf16* h = reinterpret_cast<const bf16*>(&v);
#pragma unroll
      for (int i = 0; i < 8; ++i) dyT[(grp * 8 + i) * DS + p] = h[i];
    }
    // X halo (10 x 10 px x CIT) -> xs[dx][ci][hy][j], j = hx - dx in 0..7
    for (int idx = tid; idx < 100 * (CIT / 8); idx += NT) {
      const int pos = idx / (CIT / 8), grp = idx - pos * (CIT / 8);
      const int hy = pos / 10, hx = pos - hy * 10;
      const int y = y0 + hy - 1, x = x0 + hx - 1;
      const int ci = ci0 + grp * 8;
      uint4 v = make_uint4(0, 0, 0, 0);
      if (y >= 0 && y < a.H && x >= 0 && x < a.W && ci < Cin) {
        const bf16* src;
        if (ci < a.Ca)
          src = a.A + ((size_t)(b * a.H + y) * a.W + x) * a.Ca + ci;
        else
          src = a.B + ((size_t)(b * Hh + (y >> 1)) * Wh + (x >> 1)) * a.Cb + (ci - a.Ca);
        v = *reinterpret_cast<const uint4*>(src);
      }
      const bf16* h = reinterpret_cast<const bf16*>(&v);
#pragma unroll
      for (int dx = 0; dx < 3; ++dx) {
        const int j = hx - dx;
        if (j >= 0 && j < 8) {
#pragma unroll
          for (int i = 0; i < 8; ++i) xs[(dx * CIT + grp * 8 + i) * XS + hy * 8 + j] = h[i];
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      const int prow = ks * 2 + kh;   // tile row of this lane's 8 pixels
      const bf16x8 av = as_bf8(*reinterpret_cast<const uint4*>(&dyT[(wco * 32 + row) * DS + prow * 8]));
#pragma unroll
      for (int tap = 0; tap < 9; ++tap) {
        const int dy = tap / 3, dx = tap % 3;
        const bf16x8 bv =
            as_bf8(*reinterpret_cast<const uint4*>(&xs[(dx * CIT + wci * 32 + row) * XS + (prow + dy) * 8]));
        acc[tap] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc[tap], 0, 0, 0);
      }
    }
  }
  // D[row=co][col=ci]
  const int ci = ci0 + wci * 32 + (lane & 31);
  if (ci < Cin) {
#pragma unroll
    for (int tap = 0; tap < 9; ++tap)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int co = co0 + wco * 32 + 8 * (r >> 2) + 4 * kh + (r & 3);
        a.part[(((size_t)blockIdx.z * 9 + tap) * a.Cout + co) * Cin + ci] = acc[tap][r];
      }
  }
}

// out[i] = sum_z part[z][i], z = 0..nz-1 in order
__global__ void k_reduce_splits(const float* __restrict__ part, int nz, size_t n, float* __restrict__ out) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    float s = 0.f;
    for (int z = 0; z < nz; ++z) s += part[(size_t)z * n + i];
    out[i] = s;
  }
}

// bias gradient partials: block z sums pixels [z*ppb, (z+1)*ppb) of dY for
// every channel; part[z][c]
__global__ void k_bias_partial(const bf16* __restrict__ dY, int C, size_t npix, int ppb, float* __restrict__ part) {
  __shared__ float red[256];
  const int lanes_per_c = 256 / C;   // C divides 256 (C in {32, 64, 128, 256})
  const int c = threadIdx.x % C, sub = threadIdx.x / C;
  const size_t p0 = (size_t)blockIdx.x * ppb;
  const size_t p1 = min(npix, p0 + ppb);
  float s = 0.f;
  for (size_t p = p0 + sub; p < p1; p += lanes_per_c) s += bf2f(dY[p * C + c]);
  red[threadIdx.x] = s;
  __syncthreads();
  if (sub == 0) {
    float t = 0.f;
    for (int q = 0; q < lanes_per_c; ++q) t += red[q * C + c];
    part[(size_t)blockIdx.x * C + c] = t;
  }
}

// ----------------------------------------------------------------------------
// 2x2 max pool (forward) and its gradient, which adds into dIn the pooled
// gradient at the first maximum, times (In > 0) (In is a ReLU output).
__global__ void k_maxpool(const bf16* __restrict__ in, bf16* __restrict__ out, int C, int H, int W, int batch) {
  const int Ho = H >> 1, Wo = W >> 1;
  const size_t n = (size_t)batch * Ho * Wo * C;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int c = i % C;
    size_t t = i / C;
    const int x = t % Wo;
    t /= Wo;
    const int y = t % Ho;
    const int b = t / Ho;
    const size_t r0 = ((size_t)(b * H + 2 * y) * W + 2 * x) * C + c;
    const float m = fmaxf(fmaxf(bf2f(in[r0]), bf2f(in[r0 + C])),
                          fmaxf(bf2f(in[r0 + (size_t)W * C]), bf2f(in[r0 + (size_t)W * C + C])));
    out[i] = f2bf(m);
  }
}

__global__ void k_maxpool_bwd(const bf16* __restrict__ in, const bf16* __restrict__ dout, bf16* __restrict__ din, int C,
                              int H, int W, int batch) {
  const int Ho = H >> 1, Wo = W >> 1;
  const size_t n = (size_t)batch * Ho * Wo * C;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int c = i % C;
    size_t t = i / C;
    const int x = t % Wo;
    t /= Wo;
    const int y = t % Ho;
    const int b = t / Ho;
    size_t idx[4];
    idx[0] = ((size_t)(b * H + 2 * y) * W + 2 * x) * C + c;
    idx[1] = idx[0] + C;
    idx[2] = idx[0] + (size_t)W * C;
    idx[3] = idx[2] + C;
    float v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = bf2f(in[idx[q]]);
    int am = 0;
#pragma unroll
    for (int q = 1; q < 4; ++q)
      if (v[q] > v[am]) am = q;
    const float g = bf2f(dout[i]);
    if (v[am] > 0.f) din[idx[am]] = f2bf(bf2f(din[idx[am]]) + g);
  }
}

// gradient of the nearest 2x upsampling: dB[p] = (B[p] > 0) * sum of the 4
// children of dUp (B is a ReLU output)
__global__ void k_upsample_bwd(const bf16* __restrict__ dup, const bf16* __restrict__ Bv, bf16* __restrict__ dB, int C,
                               int H, int W, int batch) {
  const int Ho = H >> 1, Wo = W >> 1;
  const size_t n = (size_t)batch * Ho * Wo * C;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int c = i % C;
    size_t t = i / C;
    const int x = t % Wo;
    t /= Wo;
    const int y = t % Ho;
    const int b = t / Ho;
    const size_t r0 = ((size_t)(b * H + 2 * y) * W + 2 * x) * C + c;
    const float s = (bf2f(dup[r0]) + bf2f(dup[r0 + C])) +
                    (bf2f(dup[r0 + (size_t)W * C]) + bf2f(dup[r0 + (size_t)W * C + C]));
    dB[i] = f2bf(bf2f(Bv[i]) > 0.f ? s : 0.f);
  }
}

// complex64 images (NHWC, C=2, FP32) -> BF16 with 16 channels (2..15 zero)
__global__ void k_pack_input(const float2* __restrict__ x, bf16* __restrict__ out, size_t npix) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < npix; i += (size_t)gridDim.x * blockDim.x) {
    const float2 v = x[i];
    uint4 q0 = make_uint4((uint32_t)f2bf(v.x) | ((uint32_t)f2bf(v.y) << 16), 0, 0, 0);
    uint4* o = reinterpret_cast<uint4*>(out + i * 16);
    o[0] = q0;
    o[1] = make_uint4(0, 0, 0, 0);
  }
}

// final 1x1 convolution (C -> 2) with the residual connection, FP32 output:
// y[p] = x[p] + bias + W f[p].  With a target, also writes dOut = 2 (y - t) / n
// and per-block partial sums of (y - t)^2.
template <int C>
__global__ __launch_bounds__(256) void k_head_fwd(const bf16* __restrict__ f, const float2* __restrict__ x,
                                                  const float* __restrict__ Wh, const float* __restrict__ bh,
                                                  float2* __restrict__ y, const float2* __restrict__ tgt,
                                                  float2* __restrict__ dout, double* __restrict__ loss_part,
                                                  size_t npix, float inv_n) {
  __shared__ float sw[2 * C + 2];
  __shared__ double red[256];
  for (int i = threadIdx.x; i < 2 * C; i += 256) sw[i] = Wh[i];
  if (threadIdx.x < 2) sw[2 * C + threadIdx.x] = bh[threadIdx.x];
  __syncthreads();
  double ls = 0.0;
  for (size_t p = blockIdx.x * (size_t)256 + threadIdx.x; p < npix; p += (size_t)gridDim.x * 256) {
    float o0 = sw[2 * C], o1 = sw[2 * C + 1];
    const uint4* fr = reinterpret_cast<const uint4*>(f + p * C);
#pragma unroll
    for (int g = 0; g < C / 8; ++g) {
      const uint4 v = fr[g];
      const bf16* h = reinterpret_cast<const bf16*>(&v);
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const float fv = bf2f(h[i]);
        o0 += sw[g * 8 + i] * fv;
        o1 += sw[C + g * 8 + i] * fv;
      }
    }
    const float2 xv = x[p];
    const float2 yv = make_float2(xv.x + o0, xv.y + o1);
    if (y) y[p] = yv;
    if (tgt) {
      const float2 tv = tgt[p];
      const float e0 = yv.x - tv.x, e1 = yv.y - tv.y;
      ls += (double)e0 * e0 + (double)e1 * e1;
      dout[p] = make_float2(2.f * e0 * inv_n, 2.f * e1 * inv_n);
    }
  }
  if (tgt) {
    red[threadIdx.x] = ls;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0];
  }
}

// head backward: df[p][c] = (f[p][c] > 0) * sum_o dOut[p][o] W[o][c] (BF16), and
// per-block partials of dW[o][c] = sum_p dOut[p][o] f[p][c], db[o] = sum_p dOut[p][o]
template <int C>
__global__ __launch_bounds__(256) void k_head_bwd(const bf16* __restrict__ f, const float2* __restrict__ dout,
                                                  const float* __restrict__ Wh, bf16* __restrict__ df,
                                                  float* __restrict__ part, size_t npix, int ppb) {
  __shared__ float red[256];
  const size_t p0 = (size_t)blockIdx.x * ppb;
  const size_t p1 = min(npix, p0 + ppb);
  // df: thread per (pixel, channel)
  for (size_t i = p0 * C + threadIdx.x; i < p1 * C; i += 256) {
    const size_t p = i / C;
    const int c = i % C;
    const float2 g = dout[p];
    const float v = g.x * Wh[c] + g.y * Wh[C + c];
    df[i] = f2bf(bf2f(f[i]) > 0.f ? v : 0.f);
  }
  // parameter-gradient partials: 2*C + 2 values; thread j handles value j
  // (j < 2C: W[o][c], o = j / C) for all pixels of the block
  for (int j0 = 0; j0 < 2 * C + 2; j0 += 256) {
    const int j = j0 + threadIdx.x;
    if (j < 2 * C + 2) {
      float s = 0.f;
      if (j < 2 * C) {
        const int o = j / C, c = j % C;
        for (size_t p = p0; p < p1; ++p) {
          const float2 g = dout[p];
          s += (o ? g.y : g.x) * bf2f(f[p * C + c]);
        }
      } else {
        const int o = j - 2 * C;
        for (size_t p = p0; p < p1; ++p) {
          const float2 g = dout[p];
          s += o ? g.y : g.x;
        }
      }
      part[(size_t)blockIdx.x * (2 * C + 2) + j] = s;
    }
  }
  (void)red;
}

// Adam (Kingma & Ba 2015) on the FP32 masters, bias-corrected step size
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       size_t n, float lr_t, float b1, float b2, float eps) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr_t * mi / (sqrtf(vi) + eps);
  }
}

// BF16 working copies of one conv layer: Wf[tap][co][ci] and the flipped,
// transposed Wb[8 - tap][ci][co]
__global__ void k_wprep(const float* __restrict__ wm, bf16* __restrict__ wf, bf16* __restrict__ wb, int Cout, int Cin) {
  const size_t n = (size_t)9 * Cout * Cin;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int ci = i % Cin;
    const size_t t = i / Cin;
    const int co = t % Cout, tap = t / Cout;
    const bf16 h = f2bf(wm[i]);
    wf[i] = h;
    wb[((size_t)(8 - tap) * Cin + ci) * Cout + co] = h;
  }
}

inline int grid_for(size_t n, int per = 256) { return (int)std::min<size_t>((n + per - 1) / per, 8192); }

}  // namespace

// ============================================================================
// host side
namespace {

struct Conv {
  int ca, cb, cout;
  int level;          // resolution level (H >> level)
  size_t w_off, b_off;
  bf16 *wf = nullptr, *wb = nullptr;
};

#define UCHK(x)                                                                          \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) {                                                              \
      cgs::set_error(std::string(#x) + " failed: " + hipGetErrorString(e_));             \
      return CGS_ERR_HIP;                                                                \
    }                                                                                    \
  } while (0)

}  // namespace

struct cgs_unet_s {
  int H = 0, W = 0, L = 4, c0 = 32, maxb = 1;
  std::vector<Conv> convs;     // enc a/b per level, then dec a/b for levels L-2..0
  size_t nparams = 0, head_w = 0, head_b = 0;
  float *params = nullptr, *grads = nullptr, *adam_m = nullptr, *adam_v = nullptr;
  long long step = 0;
  // activations (per level): a, s (enc), u, o (dec); pooled p; grads alike
  std::vector<bf16*> act_a, act_s, act_p, act_u, act_o, g_a, g_s, g_p, g_u, g_o;
  bf16* xin = nullptr;    // packed input, 16 channels
  bf16* g_up = nullptr;   // full-resolution gradient of the upsampled branch
  float2* dout = nullptr;
  float* part = nullptr;
  size_t part_n = 0;
  double* loss_part = nullptr;
  double* h_loss = nullptr;
  float2 *io_x = nullptr, *io_y = nullptr;   // staging for host I/O
  hipStream_t stream = nullptr;
  bool profile = false;
  // profile: [0] conv fwd ms, [1] conv fwd flop, [2] dgrad ms, [3] dgrad flop,
  // [4] wgrad ms, [5] wgrad flop, [6] other ms
  double stats[8] = {0};
  std::vector<hipEvent_t> ev;
};

namespace {

int level_ch(const cgs_unet_s& n, int l) { return n.c0 << l; }

size_t alloc_bytes_total = 0;

template <class T>
int dmalloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) return CGS_OK;
  hipError_t e = hipMalloc((void**)p, sizeof(T) * count);
  if (e != hipSuccess) {
    cgs::set_error("hipMalloc of " + std::to_string(sizeof(T) * count) + " bytes failed");
    return CGS_ERR_ALLOC;
  }
  alloc_bytes_total += sizeof(T) * count;
  return CGS_OK;
}

#define RCHK(x)                  \
  do {                           \
    int r_ = (x);                \
    if (r_ != CGS_OK) return r_; \
  } while (0)

struct Timer {
  cgs_unet_s& n;
  int slot;
  double flop;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  Timer(cgs_unet_s& n_, int slot_, double flop_) : n(n_), slot(slot_), flop(flop_) {
    if (!n.profile) return;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, n.stream);
  }
  ~Timer() {
    if (!n.profile) return;
    (void)hipEventRecord(e1, n.stream);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    n.stats[slot] += ms;
    if (slot < 6) n.stats[slot + 1] += flop;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
};

// forward or input-gradient convolution launch
int conv_launch(cgs_unet_s& n, const ConvArgs& a, int slot) {
  const int Cin = a.Ca + a.Cb;
  if (a.Cout % 32 || Cin % 16 || a.Ca % 8 || a.Cb % 8) {
    cgs::set_error("conv channel counts must be multiples of 32 (out) and 16 (in)");
    return CGS_ERR_UNSUPPORTED;
  }
  Timer tm(n, slot, 2.0 * 9.0 * Cin * a.Cout * (double)a.H * a.W * a.batch);
  if (a.Cout % 64 == 0 && a.W % 16 == 0 && a.H % 8 == 0) {
    dim3 g((a.W / 16) * (a.H / 8) * a.batch, a.Cout / 64);
    hipLaunchKernelGGL((k_conv3x3<8, 16, 64>), g, dim3(256), 0, n.stream, a);
  } else if (a.W % 16 == 0 && a.H % 16 == 0) {
    dim3 g((a.W / 16) * (a.H / 16) * a.batch, a.Cout / 32);
    hipLaunchKernelGGL((k_conv3x3<16, 16, 32>), g, dim3(256), 0, n.stream, a);
  } else {
    cgs::set_error("conv tile does not divide the image");
    return CGS_ERR_UNSUPPORTED;
  }
  UCHK(hipGetLastError());
  return CGS_OK;
}

int wgrad_launch(cgs_unet_s& n, const bf16* dY, const bf16* A, const bf16* B, int Ca, int Cb, int Cout, int H, int W,
                 int batch, float* dW, float* db) {
  const int Cin = Ca + Cb;
  const int ntiles = batch * (H / 8) * (W / 8);
  const int cot = (Cout % 64 == 0) ? 64 : 32;
  const int cit = (Cin % 64 == 0) ? 64 : 32;
  const int gx = Cout / cot, gy = (Cin + cit - 1) / cit;
  int nsplit = std::max(1, 1024 / (gx * gy));
  nsplit = std::min(nsplit, ntiles);
  const size_t wsz = (size_t)9 * Cout * Cin;
  while (nsplit > 1 && (size_t)nsplit * wsz > n.part_n) --nsplit;
  const int tps = (ntiles + nsplit - 1) / nsplit;
  nsplit = (ntiles + tps - 1) / tps;
  {
    Timer tm(n, 4, 2.0 * 9.0 * Cin * Cout * (double)H * W * batch);
    WgradArgs a{dY, A, B, n.part, Ca, Cb, Cout, H, W, batch, tps, ntiles};
    dim3 g(gx, gy, nsplit);
    if (cot == 64 && cit == 64)
      hipLaunchKernelGGL((k_conv3x3_wgrad<64, 64>), g, dim3(256), 0, n.stream, a);
    else if (cot == 64)
      hipLaunchKernelGGL((k_conv3x3_wgrad<64, 32>), g, dim3(128), 0, n.stream, a);
    else if (cit == 64)
      hipLaunchKernelGGL((k_conv3x3_wgrad<32, 64>), g, dim3(128), 0, n.stream, a);
    else
      hipLaunchKernelGGL((k_conv3x3_wgrad<32, 32>), g, dim3(64), 0, n.stream, a);
    UCHK(hipGetLastError());
    hipLaunchKernelGGL(k_reduce_splits, dim3(grid_for(wsz)), dim3(256), 0, n.stream, (const float*)n.part, nsplit,
                       wsz, dW);
    UCHK(hipGetLastError());
  }
  {
    Timer tm(n, 6, 0);
    const size_t npix = (size_t)batch * H * W;
    const int ppb = 1024;
    const int nb = (int)((npix + ppb - 1) / ppb);
    hipLaunchKernelGGL(k_bias_partial, dim3(nb), dim3(256), 0, n.stream, dY, Cout, npix, ppb, n.part);
    UCHK(hipGetLastError());
    hipLaunchKernelGGL(k_reduce_splits, dim3(grid_for(Cout)), dim3(256), 0, n.stream, (const float*)n.part, nb,
                       (size_t)Cout, db);
    UCHK(hipGetLastError());
  }
  return CGS_OK;
}

int prep_weights(cgs_unet_s& n) {
  for (auto& c : n.convs) {
    const int Cin = c.ca + c.cb;
    const size_t nw = (size_t)9 * c.cout * Cin;
    hipLaunchKernelGGL(k_wprep, dim3(grid_for(nw)), dim3(256), 0, n.stream, (const float*)(n.params + c.w_off), c.wf,
                       c.wb, c.cout, Cin);
    UCHK(hipGetLastError());
  }
  return CGS_OK;
}

int check_batch(const cgs_unet_s& n, int batch) {
  if (batch < 1 || batch > n.maxb) {
    cgs::set_error("batch must be 1.." + std::to_string(n.maxb));
    return CGS_ERR_ARG;
  }
  return CGS_OK;
}

// forward pass; x: device complex64 (batch, H, W); y: device or null;
// tgt: device or null (then loss partials and dOut are written)
int forward(cgs_unet_s& n, const float2* x, float2* y, const float2* tgt, int batch, double* loss) {
  const int L = n.L;
  const size_t npix = (size_t)batch * n.H * n.W;
  {
    Timer tm(n, 6, 0);
    hipLaunchKernelGGL(k_pack_input, dim3(grid_for(npix)), dim3(256), 0, n.stream, x, n.xin, npix);
    UCHK(hipGetLastError());
  }
  const bf16* in = n.xin;
  int cin = 16;
  for (int l = 0; l < L; ++l) {
    const int H = n.H >> l, W = n.W >> l, C = level_ch(n, l);
    const Conv& ca = n.convs[2 * l];
    const Conv& cb = n.convs[2 * l + 1];
    RCHK(conv_launch(n, ConvArgs{in, nullptr, ca.wf, n.params + ca.b_off, n.act_a[l], nullptr, nullptr, cin, 0, C, C, H, W,
                                 batch, 1}, 0));
    RCHK(conv_launch(n, ConvArgs{n.act_a[l], nullptr, cb.wf, n.params + cb.b_off, n.act_s[l], nullptr, nullptr, C, 0, C, C,
                                 H, W, batch, 1}, 0));
    if (l < L - 1) {
      Timer tm(n, 6, 0);
      const size_t no = (size_t)batch * (H / 2) * (W / 2) * C;
      hipLaunchKernelGGL(k_maxpool, dim3(grid_for(no)), dim3(256), 0, n.stream, (const bf16*)n.act_s[l], n.act_p[l], C,
                         H, W, batch);
      UCHK(hipGetLastError());
      in = n.act_p[l];
      cin = C;
    }
  }
  // decoder
  for (int l = L - 2; l >= 0; --l) {
    const int H = n.H >> l, W = n.W >> l, C = level_ch(n, l);
    const Conv& da = n.convs[2 * L + 2 * (L - 2 - l)];
    const Conv& db = n.convs[2 * L + 2 * (L - 2 - l) + 1];
    const bf16* below = (l + 1 == L - 1) ? n.act_s[L - 1] : n.act_o[l + 1];
    RCHK(conv_launch(n, ConvArgs{n.act_s[l], below, da.wf, n.params + da.b_off, n.act_u[l], nullptr, nullptr, C, 2 * C, C,
                                 C, H, W, batch, 1}, 0));
    RCHK(conv_launch(n, ConvArgs{n.act_u[l], nullptr, db.wf, n.params + db.b_off, n.act_o[l], nullptr, nullptr, C, 0, C, C,
                                 H, W, batch, 1}, 0));
  }
  {
    Timer tm(n, 6, 0);
    const int nb = (int)std::min<size_t>((npix + 255) / 256, 2048);
    if (n.c0 != 32) {
      cgs::set_error("head kernel is built for 32 channels");
      return CGS_ERR_UNSUPPORTED;
    }
    hipLaunchKernelGGL((k_head_fwd<32>), dim3(nb), dim3(256), 0, n.stream, (const bf16*)n.act_o[0], x,
                       (const float*)(n.params + n.head_w), (const float*)(n.params + n.head_b), y, tgt, n.dout,
                       n.loss_part, npix, (float)(1.0 / (2.0 * (double)npix)));
    UCHK(hipGetLastError());
    if (tgt && loss) {
      UCHK(hipMemcpyAsync(n.h_loss, n.loss_part, sizeof(double) * nb, hipMemcpyDeviceToHost, n.stream));
      UCHK(hipStreamSynchronize(n.stream));
      double s = 0.0;
      for (int i = 0; i < nb; ++i) s += n.h_loss[i];
      *loss = s / (2.0 * (double)npix);
    }
  }
  return CGS_OK;
}

// backward pass after forward(..., tgt, ...): fills n.grads
int backward(cgs_unet_s& n, int batch) {
  const int L = n.L;
  const size_t npix = (size_t)batch * n.H * n.W;
  {
    Timer tm(n, 6, 0);
    const int ppb = 1024;
    const int nb = (int)((npix + ppb - 1) / ppb);
    hipLaunchKernelGGL((k_head_bwd<32>), dim3(nb), dim3(256), 0, n.stream, (const bf16*)n.act_o[0],
                       (const float2*)n.dout, (const float*)(n.params + n.head_w), n.g_o[0], n.part, npix, ppb);
    UCHK(hipGetLastError());
    // head_w and head_b are adjacent (2C then 2 values)
    hipLaunchKernelGGL(k_reduce_splits, dim3(1), dim3(256), 0, n.stream, (const float*)n.part, nb,
                       (size_t)(2 * n.c0 + 2), n.grads + n.head_w);
    UCHK(hipGetLastError());
  }
  // decoder levels, l = 0 .. L-2
  for (int l = 0; l <= L - 2; ++l) {
    const int H = n.H >> l, W = n.W >> l, C = level_ch(n, l);
    const Conv& da = n.convs[2 * L + 2 * (L - 2 - l)];
    const Conv& db = n.convs[2 * L + 2 * (L - 2 - l) + 1];
    const bf16* below = (l + 1 == L - 1) ? n.act_s[L - 1] : n.act_o[l + 1];
    bf16* g_below = (l + 1 == L - 1) ? n.g_s[L - 1] : n.g_o[l + 1];
    // conv db: X = u_l, dY = g_o[l]
    RCHK(wgrad_launch(n, n.g_o[l], n.act_u[l], nullptr, C, 0, C, H, W, batch, n.grads + db.w_off, n.grads + db.b_off));
    RCHK(conv_launch(n, ConvArgs{n.g_o[l], nullptr, db.wb, nullptr, n.g_u[l], nullptr, n.act_u[l], C, 0, C, C, H, W, batch,
                                 0}, 2));
    // conv da: X = (s_l | up(below)), dY = g_u[l]
    RCHK(wgrad_launch(n, n.g_u[l], n.act_s[l], below, C, 2 * C, C, H, W, batch, n.grads + da.w_off, n.grads + da.b_off));
    RCHK(conv_launch(n, ConvArgs{n.g_u[l], nullptr, da.wb, nullptr, n.g_s[l], n.g_up, n.act_s[l], C, 0, 3 * C, C, H, W,
                                 batch, 0}, 2));
    {
      Timer tm(n, 6, 0);
      const size_t no = (size_t)batch * (H / 2) * (W / 2) * 2 * C;
      hipLaunchKernelGGL(k_upsample_bwd, dim3(grid_for(no)), dim3(256), 0, n.stream, (const bf16*)n.g_up, below,
                         g_below, 2 * C, H, W, batch);
      UCHK(hipGetLastError());
    }
  }
  // encoder levels, l = L-1 .. 0
  for (int l = L - 1; l >= 0; --l) {
    const int H = n.H >> l, W = n.W >> l, C = level_ch(n, l);
    const Conv& ca = n.convs[2 * l];
    const Conv& cb = n.convs[2 * l + 1];
    const bf16* in = l ? n.act_p[l - 1] : n.xin;
    const int cin = l ? level_ch(n, l - 1) : 16;
    RCHK(wgrad_launch(n, n.g_s[l], n.act_a[l], nullptr, C, 0, C, H, W, batch, n.grads + cb.w_off, n.grads + cb.b_off));
    RCHK(conv_launch(n, ConvArgs{n.g_s[l], nullptr, cb.wb, nullptr, n.g_a[l], nullptr, n.act_a[l], C, 0, C, C, H, W, batch,
                                 0}, 2));
    RCHK(wgrad_launch(n, n.g_a[l], in, nullptr, cin, 0, C, H, W, batch, n.grads + ca.w_off, n.grads + ca.b_off));
    if (l > 0) {
      // gradient at the pooled input (unmasked), then routed into g_s[l-1]
      RCHK(conv_launch(n, ConvArgs{n.g_a[l], nullptr, ca.wb, nullptr, n.g_p[l - 1], nullptr, nullptr, C, 0, cin, cin, H, W,
                                   batch, 0}, 2));
      Timer tm(n, 6, 0);
      const size_t no = (size_t)batch * H * W * cin;
      hipLaunchKernelGGL(k_maxpool_bwd, dim3(grid_for(no)), dim3(256), 0, n.stream, (const bf16*)n.act_s[l - 1],
                         (const bf16*)n.g_p[l - 1], n.g_s[l - 1], cin, 2 * H, 2 * W, batch);
      UCHK(hipGetLastError());
    }
  }
  return CGS_OK;
}

int upload(cgs_unet_s& n, const void* x, int batch, int device_io, const float2** dx, float2* stage) {
  if (device_io) {
    *dx = (const float2*)x;
    return CGS_OK;
  }
  UCHK(hipMemcpyAsync(stage, x, sizeof(float2) * (size_t)batch * n.H * n.W, hipMemcpyHostToDevice, n.stream));
  *dx = stage;
  return CGS_OK;
}

}  // namespace

extern "C" {

int cgs_unet_create(cgs_unet* out, int H, int W, int levels, int base_channels, int max_batch, unsigned long long seed) {
  if (!out) return CGS_ERR_ARG;
  *out = nullptr;
  if (levels < 2 || levels > 6 || base_channels != 32 || max_batch < 1) {
    cgs::set_error("unet: levels 2..6, base_channels 32, max_batch >= 1");
    return CGS_ERR_UNSUPPORTED;
  }
  const int q = 16 << (levels - 1);   // every level's side must take a 16-pixel tile
  if (H < q || W < q || H % q || W % q) {
    cgs::set_error("unet: image sides must be multiples of " + std::to_string(q));
    return CGS_ERR_UNSUPPORTED;
  }
  auto* n = new cgs_unet_s();
  n->H = H;
  n->W = W;
  n->L = levels;
  n->c0 = base_channels;
  n->maxb = max_batch;
  // layer table
  size_t off = 0;
  for (int l = 0; l < levels; ++l) {
    const int C = base_channels << l, cin = l ? (base_channels << (l - 1)) : 16;
    n->convs.push_back(Conv{cin, 0, C, l, 0, 0});
    n->convs.push_back(Conv{C, 0, C, l, 0, 0});
  }
  for (int l = levels - 2; l >= 0; --l) {
    const int C = base_channels << l;
    n->convs.push_back(Conv{C, 2 * C, C, l, 0, 0});
    n->convs.push_back(Conv{C, 0, C, l, 0, 0});
  }
  for (auto& c : n->convs) {
    c.w_off = off;
    off += (size_t)9 * c.cout * (c.ca + c.cb);
    c.b_off = off;
    off += c.cout;
  }
  n->head_w = off;
  off += 2 * (size_t)base_channels;
  n->head_b = off;
  off += 2;
  n->nparams = off;

  auto fail = [&](int rc) {
    cgs_unet_destroy(n);
    return rc;
  };
  if (hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess) {
    cgs::set_error("hipStreamCreate failed");
    delete n;
    return CGS_ERR_HIP;
  }
  int rc;
  if ((rc = dmalloc(&n->params, off)) || (rc = dmalloc(&n->grads, off)) || (rc = dmalloc(&n->adam_m, off)) ||
      (rc = dmalloc(&n->adam_v, off)))
    return fail(rc);
  for (auto& c : n->convs) {
    const size_t nw = (size_t)9 * c.cout * (c.ca + c.cb);
    if ((rc = dmalloc(&c.wf, nw)) || (rc = dmalloc(&c.wb, nw))) return fail(rc);
  }
  const size_t B = max_batch;
  n->act_a.assign(levels, nullptr);
  n->act_s.assign(levels, nullptr);
  n->act_p.assign(levels, nullptr);
  n->act_u.assign(levels, nullptr);
  n->act_o.assign(levels, nullptr);
  n->g_a.assign(levels, nullptr);
  n->g_s.assign(levels, nullptr);
  n->g_p.assign(levels, nullptr);
  n->g_u.assign(levels, nullptr);
  n->g_o.assign(levels, nullptr);
  for (int l = 0; l < levels; ++l) {
    const size_t px = B * (size_t)(H >> l) * (W >> l);
    const size_t C = (size_t)base_channels << l;
    if ((rc = dmalloc(&n->act_a[l], px * C)) || (rc = dmalloc(&n->act_s[l], px * C)) ||
        (rc = dmalloc(&n->g_a[l], px * C)) || (rc = dmalloc(&n->g_s[l], px * C)))
      return fail(rc);
    if (l < levels - 1) {
      if ((rc = dmalloc(&n->act_p[l], px / 4 * C)) || (rc = dmalloc(&n->g_p[l], px / 4 * C)) ||
          (rc = dmalloc(&n->act_u[l], px * C)) || (rc = dmalloc(&n->act_o[l], px * C)) ||
          (rc = dmalloc(&n->g_u[l], px * C)) || (rc = dmalloc(&n->g_o[l], px * C)))
        return fail(rc);
    }
  }
  const size_t px0 = B * (size_t)H * W;
  if ((rc = dmalloc(&n->xin, px0 * 16)) || (rc = dmalloc(&n->g_up, px0 * 2 * base_channels)) ||
      (rc = dmalloc(&n->dout, px0)) || (rc = dmalloc(&n->io_x, px0)) || (rc = dmalloc(&n->io_y, px0)) ||
      (rc = dmalloc(&n->loss_part, 4096)))
    return fail(rc);
  // split-K partials: the largest layer at 64 splits, at most 64 Mi floats
  size_t wmax = 0;
  for (auto& c : n->convs) wmax = std::max(wmax, (size_t)9 * c.cout * (c.ca + c.cb));
  n->part_n = std::min<size_t>(std::max<size_t>(64 * wmax, px0 * 2 / 1024 * 256 + 4096), (size_t)64 << 20);
  n->part_n = std::max(n->part_n, wmax);
  if ((rc = dmalloc(&n->part, n->part_n))) return fail(rc);
  if (hipHostMalloc((void**)&n->h_loss, sizeof(double) * 4096, 0) != hipSuccess) return fail(CGS_ERR_ALLOC);

  // He-normal initialisation (std = sqrt(2 / fan_in)), zero biases, zero head
  std::vector<float> h(off, 0.f);
  std::mt19937_64 rng(seed);
  for (auto& c : n->convs) {
    const int Cin = c.ca + c.cb;
    const int real_in = (c.level == 0 && c.ca == 16) ? 2 : Cin;   // packed input has 2 live channels
    std::normal_distribution<float> nd(0.f, std::sqrt(2.f / (9.f * real_in)));
    for (size_t i = 0; i < (size_t)9 * c.cout * Cin; ++i) h[c.w_off + i] = nd(rng);
  }
  if (hipMemcpy(n->params, h.data(), sizeof(float) * off, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(n->adam_m, 0, sizeof(float) * off) != hipSuccess ||
      hipMemset(n->adam_v, 0, sizeof(float) * off) != hipSuccess ||
      hipMemset(n->grads, 0, sizeof(float) * off) != hipSuccess) {
    cgs::set_error("unet: parameter upload failed");
    return fail(CGS_ERR_HIP);
  }
  if ((rc = prep_weights(*n)) != CGS_OK) return fail(rc);
  if (hipStreamSynchronize(n->stream) != hipSuccess) return fail(CGS_ERR_HIP);
  *out = n;
  return CGS_OK;
}

int cgs_unet_destroy(cgs_unet n) {
  if (!n) return CGS_OK;
  if (n->stream) (void)hipStreamSynchronize(n->stream);
  auto fr = [](void* p) {
    if (p) (void)hipFree(p);
  };
  fr(n->params);
  fr(n->grads);
  fr(n->adam_m);
  fr(n->adam_v);
  for (auto& c : n->convs) {
    fr(c.wf);
    fr(c.wb);
  }
  for (auto* v : {&n->act_a, &n->act_s, &n->act_p, &n->act_u, &n->act_o, &n->g_a, &n->g_s, &n->g_p, &n->g_u, &n->g_o})
    for (auto* p : *v) fr(p);
  fr(n->xin);
  fr(n->g_up);
  fr(n->dout);
  fr(n->io_x);
  fr(n->io_y);
  fr(n->part);
  fr(n->loss_part);
  if (n->h_loss) (void)hipHostFree(n->h_loss);
  if (n->stream) (void)hipStreamDestroy(n->stream);
  delete n;
  return CGS_OK;
}

int cgs_unet_num_params(cgs_unet n, long long* count) {
  if (!n || !count) return CGS_ERR_ARG;
  *count = (long long)n->nparams;
  return CGS_OK;
}

int cgs_unet_get_params(cgs_unet n, float* host) {
  if (!n || !host) return CGS_ERR_ARG;
  UCHK(hipMemcpyAsync(host, n->params, sizeof(float) * n->nparams, hipMemcpyDeviceToHost, n->stream));
  UCHK(hipStreamSynchronize(n->stream));
  return CGS_OK;
}

int cgs_unet_set_params(cgs_unet n, const float* host) {
  if (!n || !host) return CGS_ERR_ARG;
  UCHK(hipMemcpyAsync(n->params, host, sizeof(float) * n->nparams, hipMemcpyHostToDevice, n->stream));
  RCHK(prep_weights(*n));
  UCHK(hipStreamSynchronize(n->stream));
  return CGS_OK;
}

int cgs_unet_get_grads(cgs_unet n, float* host) {
  if (!n || !host) return CGS_ERR_ARG;
  UCHK(hipMemcpyAsync(host, n->grads, sizeof(float) * n->nparams, hipMemcpyDeviceToHost, n->stream));
  UCHK(hipStreamSynchronize(n->stream));
  return CGS_OK;
}

int cgs_unet_reset_optimizer(cgs_unet n) {
  if (!n) return CGS_ERR_ARG;
  UCHK(hipMemsetAsync(n->adam_m, 0, sizeof(float) * n->nparams, n->stream));
  UCHK(hipMemsetAsync(n->adam_v, 0, sizeof(float) * n->nparams, n->stream));
  UCHK(hipStreamSynchronize(n->stream));
  n->step = 0;
  return CGS_OK;
}

int cgs_unet_forward(cgs_unet n, const void* x, void* y, int batch, int device_io) {
  if (!n || !x || !y) return CGS_ERR_ARG;
  RCHK(check_batch(*n, batch));
  const float2* dx;
  RCHK(upload(*n, x, batch, device_io, &dx, n->io_x));
  float2* dy = device_io ? (float2*)y : n->io_y;
  RCHK(forward(*n, dx, dy, nullptr, batch, nullptr));
  if (!device_io)
    UCHK(hipMemcpyAsync(y, dy, sizeof(float2) * (size_t)batch * n->H * n->W, hipMemcpyDeviceToHost, n->stream));
  UCHK(hipStreamSynchronize(n->stream));
  return CGS_OK;
}

int cgs_unet_gradients(cgs_unet n, const void* x, const void* target, int batch, int device_io, double* loss) {
  if (!n || !x || !target) return CGS_ERR_ARG;
  RCHK(check_batch(*n, batch));
  const float2 *dx, *dt;
  RCHK(upload(*n, x, batch, device_io, &dx, n->io_x));
  RCHK(upload(*n, target, batch, device_io, &dt, n->io_y));
  double l = 0.0;
  RCHK(forward(*n, dx, nullptr, dt, batch, &l));
  RCHK(backward(*n, batch));
  UCHK(hipStreamSynchronize(n->stream));
  if (loss) *loss = l;
  return CGS_OK;
}

int cgs_unet_train_step(cgs_unet n, const void* x, const void* target, int batch, double lr, int device_io,
                        double* loss) {
  RCHK(cgs_unet_gradients(n, x, target, batch, device_io, loss));
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  n->step += 1;
  const double lr_t = lr * std::sqrt(1.0 - std::pow((double)b2, (double)n->step)) /
                      (1.0 - std::pow((double)b1, (double)n->step));
  {
    Timer tm(*n, 6, 0);
    hipLaunchKernelGGL(k_adam, dim3(grid_for(n->nparams)), dim3(256), 0, n->stream, n->params,
                       (const float*)n->grads, n->adam_m, n->adam_v, n->nparams, (float)lr_t, b1, b2, eps);
    UCHK(hipGetLastError());
    RCHK(prep_weights(*n));
  }
  UCHK(hipStreamSynchronize(n->stream));
  return CGS_OK;
}

int cgs_unet_profile(cgs_unet n, int enable, double* stats8) {
  if (!n) return CGS_ERR_ARG;
  if (stats8)
    for (int i = 0; i < 8; ++i) stats8[i] = n->stats[i];
  if (enable >= 0) {
    n->profile = enable != 0;
    for (double& s : n->stats) s = 0.0;
  }
  return CGS_OK;
}

}  // extern "C"
