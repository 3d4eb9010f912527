// kernels_toep.hip -- lean FFT passes of the 2D Toeplitz CG operator.
//
// Per CG iteration the Toeplitz normal operator runs three FFT passes over the
// C x 2N0 x N1 / C x 2N0 x 2N1 grid:
//   pad   p = r + beta p_old, S_c p, zero-pad, FFT along axis 0
//   conv  FFT along axis 1, times the real kernel spectrum, IFFT, crop (in place)
//   crop  IFFT along axis 0, crop, sum over 16 coils of conj(S_c) X_c
// With 32 coils and 512-point lines the generic FFT kernels issue more vector
// instructions (64-bit address arithmetic per point, run-time pad/crop tests,
// per-point exec masks) than their HBM traffic takes time.  These kernels are
// written for instruction count:
//  * image positions are not shifted by N/2: a circular convolution commutes
//    with a common shift of its input and output, so the operator is
//    unchanged, and the zero half of every padded line and the discarded half
//    of every cropped line are compile-time facts.  The first (last) DFT stage
//    is then two half-length DFTs of the even and odd bins;
//  * global accesses are raw buffer loads/stores whose per-point offsets are
//    uniform (scalar registers): one VGPR offset per thread;
//  * twiddles W_n^(a b) come from transposed LDS tables at immediate offsets,
//    constant twiddles 1, -1 and +-i are not multiplied, and the transform
//    sign is a template parameter.
// Every block takes 16 lines = 16 coils of one grid row (C % 16 == 0), so each
// global access moves 128 contiguous bytes.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdlib>

#include "cgs_internal.hpp"
#include "fft_roots.hpp"

namespace cgs {

namespace {

typedef __amdgpu_buffer_rsrc_t Rsrc;
typedef unsigned int u32x2 __attribute__((__vector_size__(2 * sizeof(unsigned int))));

// raw buffer descriptor over [base, base + bytes); word 3 = 0x00020000 selects
// the 32-bit data format of the gfx9 buffer unit
__device__ __forceinline__ Rsrc rsrc(const void* base, int64_t bytes) {
  const int nr = bytes > 0x7fffffff ? 0x7fffffff : (int)bytes;
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, nr, 0x00020000);
}
__device__ __forceinline__ float2 bld2(Rsrc r, int voff, int soff) {
  return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
__device__ __forceinline__ float bld1(Rsrc r, int voff, int soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ __forceinline__ void bst2(Rsrc r, float2 v, int voff, int soff) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, voff, soff, 0);
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {   // a * conj(b)
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

// a * exp(SG 2 pi i t / L); t is a compile-time constant after unrolling, so
// the tests fold and 1, -1, +-i cost no multiplies
template <int L, int SG>
__device__ __forceinline__ float2 rot(float2 a, int t) {
  t %= L;
  if (t == 0) return a;
  if (2 * t == L) return make_float2(-a.x, -a.y);
  if (4 * t == L) return SG > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
  if (4 * t == 3 * L) return SG > 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
  const float c = kTwL<L>.c[t];
  const float s = SG < 0 ? kTwL<L>.s[t] : -kTwL<L>.s[t];
  return make_float2(a.x * c - a.y * s, a.x * s + a.y * c);
}

// one Stockham radix-R stage (R = 2, 3 or 4), span Ns, of an N-point DFT with
// kernel exp(SG 2 pi i / N); natural order after the last stage
template <int N, int Ns, int R, int SG>
__device__ __forceinline__ void stg(const float2 (&in)[N], float2 (&out)[N]) {
#pragma unroll
  for (int jj = 0; jj < N / R; ++jj) {
    const int k = jj % Ns;
    float2 a[R];
#pragma unroll
    for (int m = 0; m < R; ++m) a[m] = in[jj + m * (N / R)];
#pragma unroll
    for (int m = 1; m < R; ++m) a[m] = rot<Ns * R, SG>(a[m], m * k);
    if constexpr (R == 2) {
      const float2 b0 = make_float2(a[0].x + a[1].x, a[0].y + a[1].y);
      const float2 b1 = make_float2(a[0].x - a[1].x, a[0].y - a[1].y);
      a[0] = b0;
      a[1] = b1;
    } else if constexpr (R == 3) {
      // w = exp(SG 2 pi i / 3) = -1/2 + SG i sqrt(3)/2
      constexpr float h = 0.86602540378443864676f * (float)SG;
      const float2 t1 = make_float2(a[1].x + a[2].x, a[1].y + a[2].y);
      const float2 t2 = make_float2(a[1].x - a[2].x, a[1].y - a[2].y);
      const float2 mm = make_float2(a[0].x - 0.5f * t1.x, a[0].y - 0.5f * t1.y);
      a[0] = make_float2(a[0].x + t1.x, a[0].y + t1.y);
      a[1] = make_float2(mm.x - h * t2.y, mm.y + h * t2.x);   // mm + (i h) t2
      a[2] = make_float2(mm.x + h * t2.y, mm.y - h * t2.x);
    } else {
      const float2 t0 = make_float2(a[0].x + a[2].x, a[0].y + a[2].y);
      const float2 t1 = make_float2(a[0].x - a[2].x, a[0].y - a[2].y);
      const float2 t2 = make_float2(a[1].x + a[3].x, a[1].y + a[3].y);
      const float2 t3 = make_float2(a[1].x - a[3].x, a[1].y - a[3].y);
      const float2 u3 = SG > 0 ? make_float2(-t3.y, t3.x) : make_float2(t3.y, -t3.x);   // (SG i) t3
      a[0] = make_float2(t0.x + t2.x, t0.y + t2.y);
      a[2] = make_float2(t0.x - t2.x, t0.y - t2.y);
      a[1] = make_float2(t1.x + u3.x, t1.y + u3.y);
      a[3] = make_float2(t1.x - u3.x, t1.y - u3.y);
    }
    const int base = (jj / Ns) * Ns * R + k;
#pragma unroll
    for (int m = 0; m < R; ++m) out[base + m * Ns] = a[m];
  }
}

// in-place N-point DFT in registers, natural order in and out
template <int N, int SG>
__device__ __forceinline__ void dft(float2 (&v)[N]) {
  float2 t[N];
  if constexpr (N == 2) {
    stg<2, 1, 2, SG>(v, t);
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = t[i];
  } else if constexpr (N == 3) {
    stg<3, 1, 3, SG>(v, t);
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = t[i];
  } else if constexpr (N == 6) {
    stg<6, 1, 2, SG>(v, t);
    stg<6, 2, 3, SG>(t, v);
  } else if constexpr (N == 12) {
    stg<12, 1, 4, SG>(v, t);
    stg<12, 4, 3, SG>(t, v);
  } else if constexpr (N == 4) {
    stg<4, 1, 4, SG>(v, t);
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = t[i];
  } else if constexpr (N == 8) {
    stg<8, 1, 2, SG>(v, t);
    stg<8, 2, 4, SG>(t, v);
  } else if constexpr (N == 16) {
    stg<16, 1, 4, SG>(v, t);
    stg<16, 4, 4, SG>(t, v);
  } else {
    static_assert(N == 32, "DFT length");
    stg<32, 1, 4, SG>(v, t);
    stg<32, 4, 4, SG>(t, v);
    stg<32, 16, 2, SG>(v, t);
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = t[i];
  }
}

// P-point DFT of v[0 .. P/2) (upper half zero): X[2r] = DFT_{P/2}(v)[r],
// X[2r+1] = DFT_{P/2}(v_k w^k)[r], w = exp(SG 2 pi i / P); all P bins on exit
template <int P, int SG>
__device__ __forceinline__ void dft_zero_top(float2 (&v)[P]) {
  constexpr int H = P / 2;
  float2 e[H], o[H];
#pragma unroll
  for (int k = 0; k < H; ++k) {
    e[k] = v[k];
    o[k] = rot<P, SG>(v[k], k);
  }
  dft<H, SG>(e);
  dft<H, SG>(o);
#pragma unroll
  for (int r = 0; r < H; ++r) {
    v[2 * r] = e[r];
    v[2 * r + 1] = o[r];
  }
}

// the first P/2 outputs of a P-point DFT: x[k] = E[k] + w^k O[k] with E, O
// the P/2-point DFTs of the even and odd inputs; v[0 .. P/2) on exit
template <int P, int SG>
__device__ __forceinline__ void dft_low_half(float2 (&v)[P]) {
  constexpr int H = P / 2;
  float2 e[H], o[H];
#pragma unroll
  for (int r = 0; r < H; ++r) {
    e[r] = v[2 * r];
    o[r] = v[2 * r + 1];
  }
  dft<H, SG>(e);
  dft<H, SG>(o);
#pragma unroll
  for (int k = 0; k < H; ++k) {
    const float2 t = rot<P, SG>(o[k], k);
    v[k] = make_float2(e[k].x + t.x, e[k].y + t.y);
  }
}

// sum over the block (every thread calls it), nw = waves per block
template <int NW>
__device__ __forceinline__ double block_sum(double v, double* red) {
  v = wave_sum_d(v);
  lds_barrier();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  lds_barrier();
  double t = 0.0;
#pragma unroll
  for (int w = 0; w < NW; ++w) t += red[w];
  return t;
}

constexpr int kXS = 16 * 16 + 8;   // conv exchange row (float2): 16 columns x 16 lines + pad
constexpr int kOutOfRange = 0x7ffffff0;   // a vector offset past every buffer's range
constexpr int kRow = 144;          // pad/crop exchange row (float2): 8 columns x 16 lines + pad

// ---------------------------------------------------------------------------
// conv: lines along axis 1 of the C x 2N0 x N1 grid (line = coil c of row l0,
// inner index c + C l0), n = 2 N1 = 16 P.  16 threads per line, thread j holds
// positions j + 16 k.  Forward: P-point DFT over k (inputs k < P/2, the image
// half), twiddle W_n^(j k1), one exchange, 16-point DFTs over j; bins
// k1 + P k2 times That; then the same steps backwards, and only positions
// j + 16 k < N1 are formed and stored.
template <int P, int SG>
__global__ void __launch_bounds__(256) k_toep_conv(ToepArgs a) {
  // P = 24 (n = 384): the second exchange round holds rows 16..23 only
  constexpr int n = 16 * P, H = P / 2, RPT = (P + 15) / 16;
  __shared__ float2 ex[16 * kXS];
  __shared__ float2 twT[16 * P];   // [j][k1] = exp(SG 2 pi i j k1 / n)
  const int tid = threadIdx.x;
  const int q = tid & 15, j = tid >> 4;
  const int64_t inner = (int64_t)a.C * 2 * a.N0;   // lines per frame
  const int64_t nb = inner / 16;
  const int f = (int)(blockIdx.x / nb);
  const int64_t q0 = ((int64_t)blockIdx.x - (int64_t)f * nb) * 16;
  if (a.fdone && a.fdone[f]) return;
  const int64_t l0 = q0 / a.C;
  float2* gb = a.out + (int64_t)f * a.N1 * inner + q0;
  const Rsrc rg = rsrc(gb, ((int64_t)a.N1 * inner - q0) * 8);
  const Rsrc rt = rsrc(a.That + ((int64_t)f * 2 * a.N0 + l0) * n, (int64_t)n * 4);
  for (int e = tid; e < 16 * P; e += 256) {
    const float2 w = a.tw[(e / P) * (e % P)];
    twT[e] = SG < 0 ? w : make_float2(w.x, -w.y);
  }
  const int vo = (int)(((int64_t)j * inner + q) * 8);
  const int so = (int)(16 * inner * 8);
  float2 v[P];
#pragma unroll
  for (int k = 0; k < H; ++k) v[k] = bld2(rg, vo, k * so);
  lds_barrier();   // twiddle table

  dft_zero_top<P, SG>(v);
#pragma unroll
  for (int k1 = 1; k1 < P; ++k1) v[k1] = cmul(v[k1], twT[j * P + k1]);
  float2 X[RPT][16];
#pragma unroll
  for (int r = 0; r < RPT; ++r) {   // round r: rows k1 = 16 r .. 16 r + 15 (< P)
    if (r) lds_barrier();
#pragma unroll
    for (int kk = 0; kk < 16; ++kk)
      if (16 * r + kk < P) ex[kk * kXS + j * 16 + q] = v[16 * r + kk];
    lds_barrier();
#pragma unroll
    for (int jj = 0; jj < 16; ++jj) X[r][jj] = ex[j * kXS + jj * 16 + q];
    dft<16, SG>(X[r]);   // rows past P (a partial round's threads j >= P - 16 r) are never read back
  }
  // thread j holds bins (j + 16 r) + P k2
  const int vt = j * 4;
#pragma unroll
  for (int r = 0; r < RPT; ++r)
    if (16 * r + j < P) {
#pragma unroll
      for (int k2 = 0; k2 < 16; ++k2) {
        const float t = bld1(rt, vt + (16 * r + P * k2) * 4, 0);
        X[r][k2] = make_float2(X[r][k2].x * t, X[r][k2].y * t);
      }
    }
#pragma unroll
  for (int r = 0; r < RPT; ++r) dft<16, -SG>(X[r]);   // over k2 -> column j'
#pragma unroll
  for (int r = 0; r < RPT; ++r) {
    lds_barrier();   // the previous round's (or the forward exchange's) reads are done
    if (16 * r + j < P) {
#pragma unroll
      for (int jj = 0; jj < 16; ++jj) ex[j * kXS + jj * 16 + q] = cmulc(X[r][jj], twT[jj * P + j + 16 * r]);
    }
    lds_barrier();
#pragma unroll
    for (int kk = 0; kk < 16; ++kk)
      if (16 * r + kk < P) v[16 * r + kk] = ex[kk * kXS + j * 16 + q];
  }
  dft_low_half<P, -SG>(v);   // over k1 -> k < H: position j + 16 k
#pragma unroll
  for (int k = 0; k < H; ++k) bst2(rg, v[k], vo, k * so);
}

// One half of k_toep_conv as a pass of its own: the middle axis (axis 1) of
// the 3D operator, lines over the inner index c + C l0 (16 per block) of each
// plane (f, i2), rows spaced inner = C 2 N0 apart.  DIR = +1: the N rows of
// `in` (the image half of the zero-padded line) -> all n = 2 N bins in `out`
// (forward, as the conv's first half); DIR = -1: the n bins of `in` -> the
// first N positions of the inverse transform in `out` (the conv's second half).
template <int P, int SG, int DIR>
__global__ void __launch_bounds__(256) k_toep_half(ToepArgs a) {
  constexpr int n = 16 * P, H = P / 2, RPT = P / 16;
  __shared__ float2 ex[16 * kXS];
  __shared__ float2 twT[16 * P];   // [j][k1] = exp(SG 2 pi i j k1 / n)
  const int tid = threadIdx.x;
  const int q = tid & 15, j = tid >> 4;
  const int64_t inner = (int64_t)a.C * 2 * a.N0;   // lines per plane
  const int64_t nb = inner / 16;
  const int64_t pl = blockIdx.x / nb;               // plane f * ppf + i2
  const int64_t q0 = ((int64_t)blockIdx.x - pl * nb) * 16;
  if (a.fdone && a.fdone[pl / a.ppf]) return;
  constexpr int64_t rin = DIR > 0 ? n / 2 : n, rout = DIR > 0 ? n : n / 2;
  const Rsrc ri = rsrc(a.in + pl * rin * inner + q0, (rin * inner - q0) * 8);
  const Rsrc ro = rsrc(a.out + pl * rout * inner + q0, (rout * inner - q0) * 8);
  for (int e = tid; e < 16 * P; e += 256) {
    const float2 w = a.tw[(e / P) * (e % P)];
    twT[e] = SG < 0 ? w : make_float2(w.x, -w.y);
  }
  const int vo = (int)(((int64_t)j * inner + q) * 8);
  const int so = (int)(16 * inner * 8);
  float2 v[P];
  float2 X[RPT][16];
  if (DIR > 0) {
#pragma unroll
    for (int k = 0; k < H; ++k) v[k] = bld2(ri, vo, k * so);
    lds_barrier();   // twiddle table
    dft_zero_top<P, SG>(v);
#pragma unroll
    for (int k1 = 1; k1 < P; ++k1) v[k1] = cmul(v[k1], twT[j * P + k1]);
#pragma unroll
    for (int r = 0; r < RPT; ++r) {   // round r: rows k1 = 16 r .. 16 r + 15
      if (r) lds_barrier();
#pragma unroll
      for (int kk = 0; kk < 16; ++kk) ex[kk * kXS + j * 16 + q] = v[16 * r + kk];
      lds_barrier();
#pragma unroll
      for (int jj = 0; jj < 16; ++jj) X[r][jj] = ex[j * kXS + jj * 16 + q];
      dft<16, SG>(X[r]);
    }
    // thread j holds bins (j + 16 r) + P k2
#pragma unroll
    for (int r = 0; r < RPT; ++r)
#pragma unroll
      for (int k2 = 0; k2 < 16; ++k2) bst2(ro, X[r][k2], vo, (int)((16 * r + P * k2) * inner * 8));
  } else {
#pragma unroll
    for (int r = 0; r < RPT; ++r)
#pragma unroll
      for (int k2 = 0; k2 < 16; ++k2) X[r][k2] = bld2(ri, vo, (int)((16 * r + P * k2) * inner * 8));
    lds_barrier();   // twiddle table
#pragma unroll
    for (int r = 0; r < RPT; ++r) dft<16, -SG>(X[r]);   // over k2 -> column j'
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
      if (r) lds_barrier();   // the previous round's reads are done
#pragma unroll
      for (int jj = 0; jj < 16; ++jj) ex[j * kXS + jj * 16 + q] = cmulc(X[r][jj], twT[jj * P + j + 16 * r]);
      lds_barrier();
#pragma unroll
      for (int kk = 0; kk < 16; ++kk) v[16 * r + kk] = ex[kk * kXS + j * 16 + q];
    }
    dft_low_half<P, -SG>(v);   // over k1 -> k < H: position j + 16 k
#pragma unroll
    for (int k = 0; k < H; ++k) bst2(ro, v[k], vo, k * so);
  }
}

// ---------------------------------------------------------------------------
// 64 threads per line, n = 64 NP (NP = 6: n = 384, NP = 8: n = 512).  A
// 1024-thread block takes 16 lines (16 coils q0 .. q0 + 15 of one grid row or
// column); thread t holds positions t + 64 k, k < NP.  Forward: an NP-point
// DFT over k, twiddle W_n^(t k1), exchange, 8-point DFT over t >> 3, twiddle
// W_64^((t & 7) m1), exchange, 8-point DFT over t & 7; thread t = 8 k1 + m1
// (t < 8 NP) then holds bins k1 + NP m1 + 8 NP m2.  Threads t >= 8 NP have no
// row in the two 8-point stages (NP = 6).
template <int NP>
struct R8 {
  static constexpr int n = 64 * NP;
  static constexpr int ACT = 8 * NP;   // threads holding a row in the 8 x 8 stages
};

// twA[t][k1] = exp(SG 2 pi i t k1 / n), twB[lo][m1] = exp(SG 2 pi i lo m1 / 64)
template <int NP, int SG>
__device__ __forceinline__ void fill_r8_tables(const float2* tw, float2* twA, float2* twB) {
  const int tid = threadIdx.x;
  if (tid < 64 * NP) {
    const float2 w = tw[(tid / NP) * (tid % NP)];
    twA[tid] = SG < 0 ? w : make_float2(w.x, -w.y);
  } else if (tid < 64 * NP + 64) {
    const int e = tid - 64 * NP;
    const float2 w = tw[NP * (e >> 3) * (e & 7)];
    twB[e] = SG < 0 ? w : make_float2(w.x, -w.y);
  }
}

// u: bins k1 of the NP-point DFT over k -> v: the inputs of the last 8-point
// DFT over t & 7 (see above)
template <int NP, int SG>
__device__ __forceinline__ void r8_forward(float2 (&u)[NP], float2 (&v)[8], float2* ex, const float2* twA,
                                           const float2* twB, int t, int q) {
  const int lo = t & 7, hi = t >> 3;
  const bool act = R8<NP>::ACT >= 64 || t < R8<NP>::ACT;
#pragma unroll
  for (int k1 = 1; k1 < NP; ++k1) u[k1] = cmul(u[k1], twA[t * NP + k1]);
#pragma unroll
  for (int k1 = 0; k1 < NP; ++k1) ex[(8 * k1 + lo) * kRow + hi * 16 + q] = u[k1];
  lds_barrier();
#pragma unroll
  for (int c = 0; c < 8; ++c) v[c] = act ? ex[t * kRow + c * 16 + q] : make_float2(0.f, 0.f);
  dft<8, SG>(v);
#pragma unroll
  for (int m1 = 1; m1 < 8; ++m1) v[m1] = cmul(v[m1], twB[lo * 8 + m1]);
  lds_barrier();
  if (act) {
#pragma unroll
    for (int m1 = 0; m1 < 8; ++m1) ex[(8 * hi + m1) * kRow + lo * 16 + q] = v[m1];
  }
  lds_barrier();
#pragma unroll
  for (int c = 0; c < 8; ++c) v[c] = act ? ex[t * kRow + c * 16 + q] : make_float2(0.f, 0.f);
  // the caller runs the last 8-point DFT over t & 7 (all bins, or m2 < 4)
}

// the same steps backwards with the conjugate kernel: v (bins m2 of thread
// 8 k1 + m1) -> u[k1] of thread t before the final NP-point DFT over k1
template <int NP, int SG>
__device__ __forceinline__ void r8_inverse(float2 (&v)[8], float2 (&u)[NP], float2* ex, const float2* twA,
                                           const float2* twB, int t, int q) {
  const int lo = t & 7, hi = t >> 3;
  const bool act = R8<NP>::ACT >= 64 || t < R8<NP>::ACT;
  dft<8, -SG>(v);   // over m2 -> j1
  lds_barrier();  // the forward exchange's reads are done
  if (act) {
#pragma unroll
    for (int j1 = 0; j1 < 8; ++j1) ex[t * kRow + j1 * 16 + q] = v[j1];
  }
  lds_barrier();
  // thread t = 8 k1 + j1 gathers over m1
#pragma unroll
  for (int m1 = 0; m1 < 8; ++m1) v[m1] = act ? ex[(8 * hi + m1) * kRow + lo * 16 + q] : make_float2(0.f, 0.f);
#pragma unroll
  for (int m1 = 1; m1 < 8; ++m1) v[m1] = cmulc(v[m1], twB[lo * 8 + m1]);
  dft<8, -SG>(v);   // over m1 -> j2
  lds_barrier();
  if (act) {
#pragma unroll
    for (int j2 = 0; j2 < 8; ++j2) ex[t * kRow + j2 * 16 + q] = v[j2];
  }
  lds_barrier();
  // thread t = j1 + 8 j2 gathers over k1
#pragma unroll
  for (int k1 = 0; k1 < NP; ++k1) u[k1] = ex[(8 * k1 + lo) * kRow + hi * 16 + q];
#pragma unroll
  for (int k1 = 1; k1 < NP; ++k1) u[k1] = cmulc(u[k1], twA[t * NP + k1]);
}

// pad: p = r + beta p_old (or x), times S_c, positions i0 = t + 64 k < N0
// (k < NP / 2), FFT along axis 0 of column (f, i1)
template <int NP, int SG>
__global__ void __launch_bounds__(1024, 8) k_toep_pad(ToepArgs a) {
  constexpr int n = R8<NP>::n, H = NP / 2;
  __shared__ float2 ex[R8<NP>::ACT * kRow];
  __shared__ float2 twA[64 * NP];
  __shared__ float2 twB[64];
  __shared__ double red[16];
  const int tid = threadIdx.x;
  const int q = tid & 15, t = tid >> 4, lo = t & 7, hi = t >> 3;
  const int nb = a.C / 16;
  const int64_t o = blockIdx.x / nb;   // f * N1 + i1
  const int q0 = (int)(blockIdx.x - o * nb) * 16;
  const int f = (int)(o / a.N1);
  const int64_t i1 = o - (int64_t)f * a.N1;
  if (a.fdone && a.fdone[f]) return;
  fill_r8_tables<NP, SG>(a.tw, twA, twB);

  const int64_t ib = (int64_t)f * a.Ntot + (int64_t)a.N0 * i1;
  const int vi = t * 8;
  const Rsrc rS = rsrc(a.S + f * a.Sfs + (int64_t)a.N0 * i1 * a.C + q0, ((int64_t)a.N0 * a.C - q0) * 8);
  const int vs = (t * a.C + q) * 8, ss = 64 * a.C * 8;
  float2 u[NP], pn[H];
  float pp = 0.f;
  if (a.pold) {   // p_new = r + beta p_old (CGLS.hpp:55-56), written by the coil-0 lanes
    const Rsrc rr = rsrc(a.r + ib, (int64_t)a.N0 * 8), rp = rsrc(a.pold + ib, (int64_t)a.N0 * 8);
    // the CG tail's ||r||^2 partials (nrpart <= kRedBlocks = 512, eight per
    // lane of wave 0) are loaded before the CG state says whether they are
    // needed, so both latencies overlap
    double rt[8];
    if (a.rpart && tid < 64) {
      const double* rpp = a.rpart + (int64_t)f * a.nrpart;
#pragma unroll
      for (int u = 0; u < 8; ++u) rt[u] = tid + 64 * u < a.nrpart ? rpp[tid + 64 * u] : 0.0;
    }
    float2 rv[H], pv[H], sv[H];
#pragma unroll
    for (int k = 0; k < H; ++k) {
      rv[k] = bld2(rr, vi, k * 512);
      pv[k] = bld2(rp, vi, k * 512);
      sv[k] = bld2(rS, vs, k * ss);
    }
    float be;
    const CgState& cs = a.cgst[f];
    if (a.rpart && cs.bpend) {
      // beta and the stop test of the iteration the CG tail just finished
      // (CGLS.hpp:53-55): wave 0 of every block sums the same ||r||^2
      // partials in the same order; the frame's first block commits the CG
      // state.  No block reads a field that block writes (the tail left
      // rn_snap and it_snap).
      __shared__ double fold_s[2];   // beta, done
      if (tid < 64) {
        double v = ((rt[0] + rt[1]) + (rt[2] + rt[3])) + ((rt[4] + rt[5]) + (rt[6] + rt[7]));
        v = wave_sum_d(v);
        if (tid == 0) {
          const double rnew = v;
          const int it = cs.it_snap + 1;
          const double rel = (a.tolkind == CGS_TOL_SQUARED_RATIO) ? rnew / cs.xn : sqrt(rnew / cs.bn);
          const int nan = (cs.nan || isnan(rel) || isnan(cs.alpha)) ? 1 : 0;
          const int done = (nan || rel <= a.tol || it >= a.maxit) ? 1 : 0;
          const double beta = rnew / cs.rn_snap;
          if ((int64_t)blockIdx.x == (int64_t)f * a.N1 * nb) {
            CgState* w = a.cgw + f;
            w->beta = beta;
            w->rn = rnew;
            w->it = it;
            w->rel = rel;
            w->nan = nan;
            w->done = done;
            if (a.hist && it < a.hcap) a.hist[(int64_t)f * a.hcap + it] = rel;
            a.fdone_w[f] = done;
          }
          fold_s[0] = beta;
          fold_s[1] = done ? 1.0 : 0.0;
        }
      }
      lds_barrier();
      if (fold_s[1] != 0.0) return;   // block-uniform
      be = (float)fold_s[0];
    } else {
      be = (float)cs.beta;
    }
    const bool c0 = (q0 + q) == 0;
#pragma unroll
    for (int k = 0; k < H; ++k) {
      pn[k] = make_float2(rv[k].x + be * pv[k].x, rv[k].y + be * pv[k].y);
      if (c0) pp += pn[k].x * pn[k].x + pn[k].y * pn[k].y;
      u[k] = cmul(pn[k], sv[k]);
    }
  } else {
    const Rsrc rx = rsrc(a.x + ib, (int64_t)a.N0 * 8);
#pragma unroll
    for (int k = 0; k < H; ++k) u[k] = cmul(bld2(rx, vi, k * 512), bld2(rS, vs, k * ss));
  }
  if (a.pp_part && q0 == 0) {   // block-uniform
    const double s = block_sum<16>(pp, red);
    if (tid == 0) a.pp_part[o] = s;
  }
  // p_new is stored after the reduction: at the merge of the two paths above
  // the compiler waits for every outstanding memory operation
  if (a.pold && q0 + q == 0) {
    const Rsrc rn = rsrc(a.pnew + ib, (int64_t)a.N0 * 8);
#pragma unroll
    for (int k = 0; k < H; ++k) bst2(rn, pn[k], vi, k * 512);
  }
  lds_barrier();   // twiddle tables

  dft_zero_top<NP, SG>(u);
  float2 v[8];
  r8_forward<NP, SG>(u, v, ex, twA, twB, t, q);
  dft<8, SG>(v);
  // bins P = hi + NP lo + 8 NP m2 of column o: G[(o * n + P) * C + q0 + q]
  if (R8<NP>::ACT >= 64 || t < R8<NP>::ACT) {
    const Rsrc ro = rsrc(a.out + o * n * a.C + q0, ((int64_t)n * a.C - q0) * 8);
    const int vb = ((hi + NP * lo) * a.C + q) * 8, sb = 8 * NP * a.C * 8;
#pragma unroll
    for (int m2 = 0; m2 < 8; ++m2) bst2(ro, v[m2], vb, m2 * sb);
  }
}

// (The group pad and crop keep two 1024-thread blocks per CU: at most 64
// VGPRs, which __launch_bounds__(1024, 8) -- eight waves per SIMD -- holds the
// compiler to.)

// crop: IFFT along axis 0 of column (f, i1), positions P = hi + NP lo +
// 8 NP m2 < N0 (m2 < 4), conj(S_c) X_c summed over the 16 coils
template <int NP, int SG>
__global__ void __launch_bounds__(1024, 8) k_toep_crop(ToepArgs a) {
  constexpr int n = R8<NP>::n;
  __shared__ float2 ex[R8<NP>::ACT * kRow];
  __shared__ float2 twA[64 * NP];
  __shared__ float2 twB[64];
  __shared__ double red[16];
  const int tid = threadIdx.x;
  const int q = tid & 15, t = tid >> 4, lo = t & 7, hi = t >> 3;
  const int nb = a.C / 16;
  const int64_t o = blockIdx.x / nb;   // f * N1 + i1
  const int g = (int)(blockIdx.x - o * nb);
  const int q0 = g * 16;
  const int f = (int)(o / a.N1);
  const int64_t i1 = o - (int64_t)f * a.N1;
  if (a.fdone && a.fdone[f]) return;
  fill_r8_tables<NP, SG>(a.tw, twA, twB);
  const Rsrc rin = rsrc(a.in + o * n * a.C + q0, ((int64_t)n * a.C - q0) * 8);
  const int vs = (t * a.C + q) * 8, ss = 64 * a.C * 8;
  float2 u[NP];
#pragma unroll
  for (int k = 0; k < NP; ++k) u[k] = bld2(rin, vs, k * ss);
  // the maps at this thread's output positions and, for <p, q>, p there: all
  // loads are issued before the transform (a buffer of 0 bytes when there is
  // no p reads zeros; positions beyond N0 of idle threads read zeros)
  const bool act = R8<NP>::ACT >= 64 || t < R8<NP>::ACT;
  const int pb = hi + NP * lo;
  const Rsrc rS = rsrc(a.S + f * a.Sfs + (int64_t)a.N0 * i1 * a.C + q0, ((int64_t)a.N0 * a.C - q0) * 8);
  const int sb = 8 * NP * a.C * 8, sp = 8 * NP * 8;
  // idle threads get a vector offset beyond every buffer's range (the range
  // check covers the vector offset, not the scalar one), so their loads read
  // zeros without touching memory
  const int vb = act ? (pb * a.C + q) * 8 : kOutOfRange, vp = act ? pb * 8 : kOutOfRange;
  const int64_t prow = (int64_t)f * a.Ntot + (int64_t)a.N0 * i1;
  const Rsrc rd = rsrc(a.pdot ? a.pdot + prow : a.in, a.pdot ? (int64_t)a.N0 * 8 : 0);
  float2 sv[4], pd[4];
#pragma unroll
  for (int m2 = 0; m2 < 4; ++m2) {
    sv[m2] = bld2(rS, vb, m2 * sb);
    pd[m2] = bld2(rd, vp, m2 * sp);
  }
  lds_barrier();   // twiddle tables

  dft<NP, SG>(u);
  float2 v[8];
  r8_forward<NP, SG>(u, v, ex, twA, twB, t, q);
  dft_low_half<8, SG>(v);   // positions hi + NP lo + 8 NP m2 < N0 (m2 < 4)
  float2 c[4];
#pragma unroll
  for (int m2 = 0; m2 < 4; ++m2) c[m2] = act ? cmulc(v[m2], sv[m2]) : make_float2(0.f, 0.f);
#pragma unroll
  for (int m2 = 0; m2 < 4; ++m2) {   // over the 16 coils (lanes q of a row)
    c[m2].x = row16_sum(c[m2].x);
    c[m2].y = row16_sum(c[m2].y);
  }
  const int64_t rowb = ((int64_t)f * nb + g) * a.Ntot + (int64_t)a.N0 * i1;
  double pq = 0.0;
  if (q == 0 && act) {
    const Rsrc ro = rsrc(a.out + rowb, (int64_t)a.N0 * 8);
#pragma unroll
    for (int m2 = 0; m2 < 4; ++m2) {
      bst2(ro, c[m2], vp, m2 * sp);
      pq += (double)pd[m2].x * c[m2].x + (double)pd[m2].y * c[m2].y;
    }
  }
  if (a.pq_part) {   // block-uniform
    const double s = block_sum<16>(pq, red);
    if (tid == 0) a.pq_part[((int64_t)f * nb + g) * a.N1 + i1] = s;
  }
}

// ---------------------------------------------------------------------------
// Row variants of pad and crop: one block per grid row (f, i1) that takes the
// row's coil groups in turn.  The loads of group g + 1 are issued before group
// g is transformed and stored, so the memory traffic of one group overlaps
// the arithmetic of the next; with one group per block the resident blocks run
// load, compute and store in lock-step rounds and the phases do not overlap.
// The pad forms p = r + beta p_old once per row; the crop sums conj(S_c) X_c
// over all C coils, so it writes the whole coil sum (one partial image).

template <int NP, int SG, int G>
__global__ void __launch_bounds__(1024) k_toep_pad_row(ToepArgs a) {
  constexpr int n = R8<NP>::n, H = NP / 2;
  __shared__ float2 ex[R8<NP>::ACT * kRow];
  __shared__ float2 twA[64 * NP];
  __shared__ float2 twB[64];
  __shared__ double red[16];
  const int tid = threadIdx.x;
  const int q = tid & 15, t = tid >> 4, lo = t & 7, hi = t >> 3;
  const int64_t o = blockIdx.x;   // f * N1 + i1
  const int f = (int)(o / a.N1);
  const int64_t i1 = o - (int64_t)f * a.N1;
  if (a.fdone && a.fdone[f]) return;
  // the folded tail's per-row partials (<p, q>, <r, q>, |q|^2, ||r||^2 of rows
  // tid + 64 j), loaded first so that their latency overlaps the loads below
  double tfp[4][4];
  if (a.tailfold && tid < 64) {
    const int64_t fo = (int64_t)f * a.nrpart;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int u = tid + 64 * j;
      const bool in = u < a.nrpart;
      tfp[0][j] = in ? a.fpq[fo + u] : 0.0;
      tfp[1][j] = in ? a.frq[fo + u] : 0.0;
      tfp[2][j] = in ? a.fqq[fo + u] : 0.0;
      tfp[3][j] = in ? a.frr[fo + u] : 0.0;
    }
  }
  fill_r8_tables<NP, SG>(a.tw, twA, twB);

  const int64_t ib = (int64_t)f * a.Ntot + (int64_t)a.N0 * i1;
  const int vi = t * 8;
  // group g's 16 coils sit 128 bytes after group g - 1's
  const Rsrc rS = rsrc(a.S + f * a.Sfs + (int64_t)a.N0 * i1 * a.C, (int64_t)a.N0 * a.C * 8);
  const int vs = (t * a.C + q) * 8, ss = 64 * a.C * 8;
  float2 sv[H], sn[H], xv[H];
#pragma unroll
  for (int k = 0; k < H; ++k) sv[k] = bld2(rS, vs, k * ss);
  float pp = 0.f;
  bool fold = false;   // folded CG tail: x, r and p_new leave through the LDS staging below
  if (a.pold) {   // p_new = r + beta p_old (CGLS.hpp:55-56), written by the q = 0 lanes
    const Rsrc rr = rsrc(a.r + ib, (int64_t)a.N0 * 8), rp = rsrc(a.pold + ib, (int64_t)a.N0 * 8);
    double rt[8];
    if (a.rpart && tid < 64) {
      const double* rpp = a.rpart + (int64_t)f * a.nrpart;
#pragma unroll
      for (int u = 0; u < 8; ++u) rt[u] = tid + 64 * u < a.nrpart ? rpp[tid + 64 * u] : 0.0;
    }
    float2 rv[H], pv[H], qrow[H], xrow[H];
    if (a.tailfold) {   // block-uniform
      // r, p_old, q and x of the row: one coalesced load per thread into LDS
      // (the exchange buffer, unused until the FFT), then each thread takes
      // its positions t + 64 k
      constexpr int N0 = 32 * NP;
      for (int e = tid; e < 4 * N0; e += 1024) {
        const int arr = e / N0, pos = e - arr * N0;
        const float2* src = arr == 0 ? a.r : arr == 1 ? a.pold : arr == 2 ? a.qv : a.xs;
        ex[e] = src[ib + pos];
      }
      lds_barrier();
#pragma unroll
      for (int k = 0; k < H; ++k) {
        rv[k] = ex[t + 64 * k];
        pv[k] = ex[N0 + t + 64 * k];
        qrow[k] = ex[2 * N0 + t + 64 * k];
        xrow[k] = ex[3 * N0 + t + 64 * k];
      }
    } else {
#pragma unroll
      for (int k = 0; k < H; ++k) {
        rv[k] = bld2(rr, vi, k * 512);
        pv[k] = bld2(rp, vi, k * 512);
      }
    }
    float be;
    const CgState& cs = a.cgst[f];
    if (a.tailfold && cs.bpend) {
      // the previous iteration's CG tail: alpha from the crop's <p, q>,
      // ||r_new||^2 = R - 2 alpha <r, q> + alpha^2 |q|^2 (r_new = r - alpha q),
      // beta and the stop test; x and r of this row updated here
      __shared__ double tf_s[3];
      if (tid < 64) {
        const int64_t fo = (int64_t)f * a.nrpart;
        double sp = 0.0, sr = 0.0, sq = 0.0, R = 0.0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {   // the same order in every block
          sp += tfp[0][j];
          sr += tfp[1][j];
          sq += tfp[2][j];
          R += tfp[3][j];
        }
        for (int u = tid + 256; u < a.nrpart; u += 64) {
          sp += a.fpq[fo + u];
          sr += a.frq[fo + u];
          sq += a.fqq[fo + u];
          R += a.frr[fo + u];
        }
        sp = wave_sum_d(sp);
        sr = wave_sum_d(sr);
        sq = wave_sum_d(sq);
        R = wave_sum_d(R);
        if (tid == 0) {
          // R: ||r||^2 of the stored residual, so that rounding in the
          // recurrence does not accumulate over iterations
          const double al = R / sp;   // CGLS.hpp:50
          const double g = fmax(R - 2.0 * al * sr + al * al * sq, 0.0);
          const int it = cs.it_snap + 1;
          const double rel = (a.tolkind == CGS_TOL_SQUARED_RATIO) ? g / cs.xn : sqrt(g / cs.bn);
          const int nan = (cs.nan || isnan(rel) || isnan(al)) ? 1 : 0;
          const int done = (nan || rel <= a.tol || it >= a.maxit) ? 1 : 0;
          const double beta = g / R;
          if (i1 == 0) {   // the frame's first block commits (the crop passes `done` on)
            CgState* w = a.cgw + f;
            w->pq = sp;
            w->alpha = al;
            w->beta = beta;
            w->rn = g;
            w->it = it;
            w->rel = rel;
            w->nan = nan;
            w->done = done;
            if (a.hist && it < a.hcap) a.hist[(int64_t)f * a.hcap + it] = rel;
          }
          tf_s[0] = al;
          tf_s[1] = beta;
          tf_s[2] = done ? 1.0 : 0.0;
        }
      }
      lds_barrier();
      const float al = (float)tf_s[0];
#pragma unroll
      for (int k = 0; k < H; ++k) {
        xrow[k] = make_float2(xrow[k].x + al * pv[k].x, xrow[k].y + al * pv[k].y);
        rv[k] = make_float2(rv[k].x - al * qrow[k].x, rv[k].y - al * qrow[k].y);
      }
      if (tf_s[2] != 0.0) {   // block-uniform: the frame stops here with these x and r
        if (q == 0) {
          const Rsrc wx = rsrc(a.xs + ib, (int64_t)a.N0 * 8), wr = rsrc(a.rs + ib, (int64_t)a.N0 * 8);
#pragma unroll
          for (int k = 0; k < H; ++k) {
            bst2(wx, xrow[k], vi, k * 512);
            bst2(wr, rv[k], vi, k * 512);
          }
        }
        return;
      }
      be = (float)tf_s[1];
      fold = true;
    } else if (a.rpart && cs.bpend) {
      // beta and the stop test of the finished iteration, as in k_toep_pad
      __shared__ double fold_s[2];
      if (tid < 64) {
        double v = ((rt[0] + rt[1]) + (rt[2] + rt[3])) + ((rt[4] + rt[5]) + (rt[6] + rt[7]));
        v = wave_sum_d(v);
        if (tid == 0) {
          const double rnew = v;
          const int it = cs.it_snap + 1;
          const double rel = (a.tolkind == CGS_TOL_SQUARED_RATIO) ? rnew / cs.xn : sqrt(rnew / cs.bn);
          const int nan = (cs.nan || isnan(rel) || isnan(cs.alpha)) ? 1 : 0;
          const int done = (nan || rel <= a.tol || it >= a.maxit) ? 1 : 0;
          const double beta = rnew / cs.rn_snap;
          if (i1 == 0) {   // the frame's first block commits
            CgState* w = a.cgw + f;
            w->beta = beta;
            w->rn = rnew;
            w->it = it;
            w->rel = rel;
            w->nan = nan;
            w->done = done;
            if (a.hist && it < a.hcap) a.hist[(int64_t)f * a.hcap + it] = rel;
            a.fdone_w[f] = done;
          }
          fold_s[0] = beta;
          fold_s[1] = done ? 1.0 : 0.0;
        }
      }
      lds_barrier();
      if (fold_s[1] != 0.0) return;   // block-uniform
      be = (float)fold_s[0];
    } else {
      be = (float)cs.beta;
    }
#pragma unroll
    for (int k = 0; k < H; ++k) {
      xv[k] = make_float2(rv[k].x + be * pv[k].x, rv[k].y + be * pv[k].y);
      if (q == 0) pp += xv[k].x * xv[k].x + xv[k].y * xv[k].y;
    }
    if (fold) {   // block-uniform
      // x, r and p_new of the row: a thread holds positions t + 64 k, so
      // they go through LDS (the exchange buffer, unused until the FFT) and
      // leave in runs of 64 consecutive positions per wave
      constexpr int N0 = 32 * NP;
      if (q == 0) {
#pragma unroll
        for (int k = 0; k < H; ++k) {
          ex[t + 64 * k] = xrow[k];
          ex[N0 + t + 64 * k] = rv[k];
          ex[2 * N0 + t + 64 * k] = xv[k];
        }
      }
      lds_barrier();
      for (int e = tid; e < 3 * N0; e += 1024) {
        const int arr = e / N0, pos = e - arr * N0;
        float2* dst = arr == 0 ? a.xs : arr == 1 ? a.rs : a.pnew;
        dst[ib + pos] = ex[e];
      }
    }
  } else {
    const Rsrc rx = rsrc(a.x + ib, (int64_t)a.N0 * 8);
#pragma unroll
    for (int k = 0; k < H; ++k) xv[k] = bld2(rx, vi, k * 512);
  }
  if (a.pp_part) {   // block-uniform
    const double s = block_sum<16>(pp, red);
    if (tid == 0) a.pp_part[o] = s;
  }
  // p_new is stored after the reduction: at the merge of the two paths above
  // the compiler waits for every outstanding memory operation
  if (a.pold && q == 0 && !fold) {
    const Rsrc rn = rsrc(a.pnew + ib, (int64_t)a.N0 * 8);
#pragma unroll
    for (int k = 0; k < H; ++k) bst2(rn, xv[k], vi, k * 512);
  }
  lds_barrier();   // twiddle tables

  const bool act = R8<NP>::ACT >= 64 || t < R8<NP>::ACT;
  const Rsrc ro = rsrc(a.out + o * n * a.C, (int64_t)n * a.C * 8);
  const int vb = ((hi + NP * lo) * a.C + q) * 8, sb = 8 * NP * a.C * 8;
#pragma unroll
  for (int g = 0; g < G; ++g) {
    if (g + 1 < G) {
#pragma unroll
      for (int k = 0; k < H; ++k) sn[k] = bld2(rS, vs, k * ss + (g + 1) * 128);
    }
    float2 u[NP];
#pragma unroll
    for (int k = 0; k < H; ++k) u[k] = cmul(xv[k], sv[k]);
    if (g > 0) lds_barrier();   // the previous group's exchange reads are done
    dft_zero_top<NP, SG>(u);
    float2 v[8];
    r8_forward<NP, SG>(u, v, ex, twA, twB, t, q);
    dft<8, SG>(v);
    if (act) {
#pragma unroll
      for (int m2 = 0; m2 < 8; ++m2) bst2(ro, v[m2], vb, m2 * sb + g * 128);
    }
#pragma unroll
    for (int k = 0; k < H; ++k) sv[k] = sn[k];
  }
}

template <int NP, int SG, int G>
__global__ void __launch_bounds__(1024) k_toep_crop_row(ToepArgs a) {
  constexpr int n = R8<NP>::n;
  __shared__ float2 ex[R8<NP>::ACT * kRow];
  __shared__ float2 twA[64 * NP];
  __shared__ float2 twB[64];
  __shared__ double red[16];
  const int tid = threadIdx.x;
  const int q = tid & 15, t = tid >> 4, lo = t & 7, hi = t >> 3;
  const int64_t o = blockIdx.x;   // f * N1 + i1
  const int f = (int)(o / a.N1);
  const int64_t i1 = o - (int64_t)f * a.N1;
  if (a.fdone && a.fdone[f]) return;
  if (a.tailfold && a.cgst[f].done) {   // the pad stopped this frame: pass the flag on
    if (i1 == 0 && tid == 0) a.fdone_w[f] = 1;
    return;
  }
  fill_r8_tables<NP, SG>(a.tw, twA, twB);
  const Rsrc rin = rsrc(a.in + o * n * a.C, (int64_t)n * a.C * 8);
  const int vs = (t * a.C + q) * 8, ss = 64 * a.C * 8;
  float2 u[NP], un[NP];
#pragma unroll
  for (int k = 0; k < NP; ++k) u[k] = bld2(rin, vs, k * ss);
  const bool act = R8<NP>::ACT >= 64 || t < R8<NP>::ACT;
  const int pb = hi + NP * lo;
  const Rsrc rS = rsrc(a.S + f * a.Sfs + (int64_t)a.N0 * i1 * a.C, (int64_t)a.N0 * a.C * 8);
  const int sb = 8 * NP * a.C * 8, sp = 8 * NP * 8;
  const int vb = act ? (pb * a.C + q) * 8 : kOutOfRange, vp = act ? pb * 8 : kOutOfRange;
  const int64_t rowb = (int64_t)f * a.Ntot + (int64_t)a.N0 * i1;
  // p at this thread's output positions for <p, q>, loaded ahead (a buffer
  // of 0 bytes when there is no p reads zeros)
  const Rsrc rd = rsrc(a.pdot ? a.pdot + rowb : a.in, a.pdot ? (int64_t)a.N0 * 8 : 0);
  float2 pd[4];
#pragma unroll
  for (int m2 = 0; m2 < 4; ++m2) pd[m2] = bld2(rd, vp, m2 * sp);
  // r at the same positions, for <r, q> and ||r||^2 when the CG tail is folded (zeros otherwise)
  const Rsrc rrd = rsrc(a.rdot ? a.rdot + rowb : a.in, a.rdot ? (int64_t)a.N0 * 8 : 0);
  float2 rdv[4];
#pragma unroll
  for (int m2 = 0; m2 < 4; ++m2) rdv[m2] = bld2(rrd, vp, m2 * sp);
  float2 c[4];
#pragma unroll
  for (int m2 = 0; m2 < 4; ++m2) c[m2] = make_float2(0.f, 0.f);
  lds_barrier();   // twiddle tables
#pragma unroll
  for (int g = 0; g < G; ++g) {
    if (g + 1 < G) {
#pragma unroll
      for (int k = 0; k < NP; ++k) un[k] = bld2(rin, vs, k * ss + (g + 1) * 128);
    }
    float2 sv[4];
#pragma unroll
    for (int m2 = 0; m2 < 4; ++m2) sv[m2] = bld2(rS, vb, m2 * sb + g * 128);
    if (g > 0) lds_barrier();   // the previous group's exchange reads are done
    dft<NP, SG>(u);
    float2 v[8];
    r8_forward<NP, SG>(u, v, ex, twA, twB, t, q);
    dft_low_half<8, SG>(v);   // positions hi + NP lo + 8 NP m2 < N0 (m2 < 4)
#pragma unroll
    for (int m2 = 0; m2 < 4; ++m2) {
      const float2 w = act ? cmulc(v[m2], sv[m2]) : make_float2(0.f, 0.f);
      c[m2].x += w.x;
      c[m2].y += w.y;
    }
#pragma unroll
    for (int k = 0; k < NP; ++k) u[k] = un[k];
  }
#pragma unroll
  for (int m2 = 0; m2 < 4; ++m2) {   // over the 16 coils (lanes q of a row)
    c[m2].x = row16_sum(c[m2].x);
    c[m2].y = row16_sum(c[m2].y);
  }
  double pq = 0.0, rq = 0.0, qq = 0.0, rr = 0.0;
  if (q == 0 && act) {
    const Rsrc ro = rsrc(a.out + rowb, (int64_t)a.N0 * 8);
#pragma unroll
    for (int m2 = 0; m2 < 4; ++m2) {
      float2 qv = c[m2];
      if (a.tailfold) {   // q = (E^H W E) p + lambda p
        qv.x += a.lam * pd[m2].x;
        qv.y += a.lam * pd[m2].y;
        rq += (double)rdv[m2].x * qv.x + (double)rdv[m2].y * qv.y;
        qq += (double)qv.x * qv.x + (double)qv.y * qv.y;
        rr += (double)rdv[m2].x * rdv[m2].x + (double)rdv[m2].y * rdv[m2].y;
      }
      bst2(ro, qv, vp, m2 * sp);
      pq += (double)pd[m2].x * qv.x + (double)pd[m2].y * qv.y;
    }
  }
  if (a.tailfold) {   // block-uniform: the four sums over one exchange
    __shared__ double red4[4 * 16];
    pq = wave_sum_d(pq);
    rq = wave_sum_d(rq);
    qq = wave_sum_d(qq);
    rr = wave_sum_d(rr);
    if ((tid & 63) == 0) {
      red4[(tid >> 6)] = pq;
      red4[16 + (tid >> 6)] = rq;
      red4[32 + (tid >> 6)] = qq;
      red4[48 + (tid >> 6)] = rr;
    }
    lds_barrier();
    if (tid < 4) {
      double v = 0.0;
#pragma unroll
      for (int w = 0; w < 16; ++w) v += red4[16 * tid + w];
      double* dst = tid == 0 ? a.pq_part : tid == 1 ? a.rq_part : tid == 2 ? a.qq_part : a.rr_part;
      dst[o] = v;
    }
  } else if (a.pq_part) {   // block-uniform
    const double s = block_sum<16>(pq, red);
    if (tid == 0) a.pq_part[o] = s;
  }
  if (a.tailfold) {
    // the iteration's rn and it for the next pad, whose first block rewrites them
    if (i1 == 0 && tid == 0) {
      CgState* w = a.cgw + f;
      w->rn_snap = w->rn;
      w->it_snap = w->it;
      w->bpend = 1;
    }
  }
}

// conv on 64 threads per line (n = 2 N1 = 64 NP): lines along axis 1 as in
// k_toep_conv, positions i1 = t + 64 k < N1 (k < NP / 2) in and out
template <int NP, int SG>
__global__ void __launch_bounds__(1024) k_toep_conv_r8(ToepArgs a) {
  constexpr int n = R8<NP>::n, H = NP / 2;
  __shared__ float2 ex[R8<NP>::ACT * kRow];
  __shared__ float2 twA[64 * NP];
  __shared__ float2 twB[64];
  const int tid = threadIdx.x;
  const int q = tid & 15, t = tid >> 4, lo = t & 7, hi = t >> 3;
  const int64_t inner = (int64_t)a.C * 2 * a.N0;   // lines per frame
  const int64_t nb = inner / 16;
  const int f = (int)(blockIdx.x / nb);
  const int64_t q0 = ((int64_t)blockIdx.x - (int64_t)f * nb) * 16;
  if (a.fdone && a.fdone[f]) return;
  const int64_t l0 = q0 / a.C;
  fill_r8_tables<NP, SG>(a.tw, twA, twB);
  float2* gb = a.out + (int64_t)f * a.N1 * inner + q0;
  const Rsrc rg = rsrc(gb, ((int64_t)a.N1 * inner - q0) * 8);
  const int vo = (int)(((int64_t)t * inner + q) * 8);
  const int so = (int)(64 * inner * 8);
  float2 u[NP];
#pragma unroll
  for (int k = 0; k < H; ++k) u[k] = bld2(rg, vo, k * so);
  lds_barrier();   // twiddle tables

  dft_zero_top<NP, SG>(u);
  float2 v[8];
  r8_forward<NP, SG>(u, v, ex, twA, twB, t, q);
  dft<8, SG>(v);
  if (R8<NP>::ACT >= 64 || t < R8<NP>::ACT) {
    const Rsrc rt = rsrc(a.That + ((int64_t)f * 2 * a.N0 + l0) * n, (int64_t)n * 4);
    const int vt = (hi + NP * lo) * 4;
#pragma unroll
    for (int m2 = 0; m2 < 8; ++m2) {
      const float w = bld1(rt, vt + m2 * 8 * NP * 4, 0);
      v[m2] = make_float2(v[m2].x * w, v[m2].y * w);
    }
  }
  r8_inverse<NP, SG>(v, u, ex, twA, twB, t, q);
  dft_low_half<NP, -SG>(u);   // positions t + 64 k, k < H
#pragma unroll
  for (int k = 0; k < H; ++k) bst2(rg, u[k], vo, k * so);
}

}  // namespace

// C % 16 == 0 and every per-frame byte offset below 2^31: pad/crop at 2 N0 =
// 256, 384 or 512 (64 threads per line), the conv at 2 N1 = 256 or 512
// (16 threads per line) or 384 (64 threads per line)
bool toep_pad_crop_supported(int N0, int N1, int C) {
  return (N0 == 128 || N0 == 192 || N0 == 256) && C % 16 == 0 && N1 > 0 &&
         (int64_t)2 * N0 * C * 8 < (int64_t)1 << 31;
}
// the 3D middle-axis passes: n = 2 N1 = 256 or 512, every per-plane offset
// below 2^31
bool toep_half_supported(int N0, int N1, int C) {
  return (N1 == 128 || N1 == 256) && C % 16 == 0 && (int64_t)2 * N1 * C * 2 * N0 * 8 < (int64_t)1 << 31;
}
bool toep_conv_supported(int N0, int N1, int C) {
  return (N1 == 128 || N1 == 192 || N1 == 256) && C % 16 == 0 && N0 > 0 &&
         (int64_t)N1 * C * 2 * N0 * 8 < (int64_t)1 << 31;
}

// blocks resident per CU: unused dynamic LDS caps it at `occ` (0: no cap), so
// that blocks of a later round start while others compute or store
static size_t occ_pad_lds(const char* env, size_t static_lds) {
  const char* e = std::getenv(env);
  const int occ = e ? std::atoi(e) : 0;
  if (occ <= 0) return 0;
  const size_t per = (size_t)163840 / (size_t)occ;
  return per > static_lds + 256 ? per - static_lds - 256 : 0;
}

// NP = N0 / 32 points per thread (n = 2 N0 = 64 NP): 4, 6 or 8
template <int NP>
static void launch_pad_np(const ToepArgs& a, hipStream_t s) {
  if (a.row) {
    const dim3 grid((unsigned)((int64_t)a.F * a.N1));
    const int G = a.C / 16;
    if (G == 2) CGS_LAUNCH((k_toep_pad_row<NP, -1, 2>), grid, dim3(1024), 0, s, a);
    else if (G == 3) CGS_LAUNCH((k_toep_pad_row<NP, -1, 3>), grid, dim3(1024), 0, s, a);
    else CGS_LAUNCH((k_toep_pad_row<NP, -1, 4>), grid, dim3(1024), 0, s, a);
    return;
  }
  const dim3 grid((unsigned)((int64_t)a.F * a.N1 * (a.C / 16)));
  static const size_t extra = NP == 8 ? occ_pad_lds("CGS_TOEP_OCC_PC", 78464) : 0;
  CGS_LAUNCH((k_toep_pad<NP, -1>), grid, dim3(1024), extra, s, a);
}

void launch_toep_pad(const ToepArgs& a, hipStream_t s) {
  if (a.N0 == 256) launch_pad_np<8>(a, s);
  else if (a.N0 == 192) launch_pad_np<6>(a, s);
  else launch_pad_np<4>(a, s);
}

void launch_toep_conv(const ToepArgs& a, hipStream_t s) {
  const int64_t lines = (int64_t)a.F * a.C * 2 * a.N0;
  static const size_t extra = occ_pad_lds("CGS_TOEP_OCC_CONV", 37888);
  if (a.N1 == 256) CGS_LAUNCH((k_toep_conv<32, -1>), dim3((unsigned)(lines / 16)), dim3(256), extra, s, a);
  else if (a.N1 == 128) CGS_LAUNCH((k_toep_conv<16, -1>), dim3((unsigned)(lines / 16)), dim3(256), extra, s, a);
  else {   // N1 = 192: 16 threads per line (P = 24) unless CGS_TOEP_CONV24=0 (64 threads per line)
    static const bool c24 = [] { const char* e = std::getenv("CGS_TOEP_CONV24"); return !(e && std::atoi(e) == 0); }();
    if (c24) CGS_LAUNCH((k_toep_conv<24, -1>), dim3((unsigned)(lines / 16)), dim3(256), 0, s, a);
    else CGS_LAUNCH((k_toep_conv_r8<6, -1>), dim3((unsigned)(lines / 16)), dim3(1024), 0, s, a);
  }
}

template <int NP>
static void launch_crop_np(const ToepArgs& a, hipStream_t s) {
  if (a.row) {
    const dim3 grid((unsigned)((int64_t)a.F * a.N1));
    const int G = a.C / 16;
    if (G == 2) CGS_LAUNCH((k_toep_crop_row<NP, 1, 2>), grid, dim3(1024), 0, s, a);
    else if (G == 3) CGS_LAUNCH((k_toep_crop_row<NP, 1, 3>), grid, dim3(1024), 0, s, a);
    else CGS_LAUNCH((k_toep_crop_row<NP, 1, 4>), grid, dim3(1024), 0, s, a);
    return;
  }
  const dim3 grid((unsigned)((int64_t)a.F * a.N1 * (a.C / 16)));
  static const size_t extra = NP == 8 ? occ_pad_lds("CGS_TOEP_OCC_PC", 78464) : 0;
  CGS_LAUNCH((k_toep_crop<NP, 1>), grid, dim3(1024), extra, s, a);
}

void launch_toep_crop(const ToepArgs& a, hipStream_t s) {
  if (a.N0 == 256) launch_crop_np<8>(a, s);
  else if (a.N0 == 192) launch_crop_np<6>(a, s);
  else launch_crop_np<4>(a, s);
}

// the middle-axis half passes of the 3D operator: lines of n = 2 N1 = 256 or
// 512 points over the planes (f, i2)
void launch_toep_half(const ToepArgs& a, int dir, hipStream_t s) {
  const int64_t lines = (int64_t)a.F * a.ppf * a.C * 2 * a.N0;
  const dim3 grid((unsigned)(lines / 16));
  if (dir > 0) {
    if (a.N1 == 256) CGS_LAUNCH((k_toep_half<32, -1, 1>), grid, dim3(256), 0, s, a);
    else CGS_LAUNCH((k_toep_half<16, -1, 1>), grid, dim3(256), 0, s, a);
  } else {
    if (a.N1 == 256) CGS_LAUNCH((k_toep_half<32, -1, -1>), grid, dim3(256), 0, s, a);
    else CGS_LAUNCH((k_toep_half<16, -1, -1>), grid, dim3(256), 0, s, a);
  }
}

}  // namespace cgs
