// kernels_fft_reg.hip -- register-resident FFT passes for line lengths 256 and
// 512 (the oversampled / Toeplitz grids of the 2D workloads).
//
// Same job and arguments as k_fft (kernels_fft.hip): a pass views the array as
// [outer][n][inner] with inner contiguous and transforms the n-point lines.  A
// 256-thread block takes 16 adjacent lines (16 consecutive inner indices, so
// every global access moves 128 contiguous bytes); 16 threads share a line and
// each holds N1 = n / 16 of its points in registers.  With n = 16 * N1 and
// positions p = j + 16 k (thread j, register k):
//   X[k1 + N1 k2] = sum_j W_16^(j k2) W_n^(j k1) [ sum_k x[j + 16 k] W_N1^(k k1) ]
// so a pass is an N1-point FFT in registers, a twiddle, ONE exchange through
// LDS, and N1/16 16-point FFTs in registers.  The Toeplitz convolution pass
// multiplies the spectrum in that layout and runs the same steps backwards
// (two exchanges in all).  Compared with staging whole lines in LDS and
// running log_R(n) LDS stages per line, LDS traffic drops from ~4 reads and
// writes per point to one of each, and all N1 loads per thread are in flight
// together.
#include <hip/hip_runtime.h>

#include <cstdlib>

#include "cgs_internal.hpp"
#include "fft_roots.hpp"

namespace cgs {

namespace {

__device__ __forceinline__ float2 cmul2(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// one Stockham radix-R stage (R = 2 or 4) of an N-point DFT held in registers;
// span Ns; sign s = -1 forward, +1 inverse; every index is a compile-time
// constant after unrolling, so the arrays stay in VGPRs
template <int N, int Ns, int R>
__device__ __forceinline__ void reg_stage(const float2 (&in)[N], float2 (&out)[N], float s) {
#pragma unroll
  for (int jj = 0; jj < N / R; ++jj) {
    const int k = jj % Ns;
    float2 a[R];
#pragma unroll
    for (int m = 0; m < R; ++m) a[m] = in[jj + m * (N / R)];
    if (Ns > 1) {
#pragma unroll
      for (int m = 1; m < R; ++m) {
        const int t = (m * k) % (Ns * R);
        const float2 w = make_float2(kTwL<Ns * R>.c[t], s < 0 ? kTwL<Ns * R>.s[t] : -kTwL<Ns * R>.s[t]);
        a[m] = cmul2(a[m], w);
      }
    }
    if (R == 2) {
      const float2 b0 = make_float2(a[0].x + a[1].x, a[0].y + a[1].y);
      const float2 b1 = make_float2(a[0].x - a[1].x, a[0].y - a[1].y);
      a[0] = b0;
      a[1] = b1;
    } else if (R == 3) {
      // w = exp(s 2 pi i / 3) = -1/2 + s i sqrt(3)/2
      const float h = 0.86602540378443864676f * s;
      const float2 t1 = make_float2(a[1].x + a[2].x, a[1].y + a[2].y);
      const float2 t2 = make_float2(a[1].x - a[2].x, a[1].y - a[2].y);
      const float2 b0 = make_float2(a[0].x + t1.x, a[0].y + t1.y);
      const float2 m = make_float2(a[0].x - 0.5f * t1.x, a[0].y - 0.5f * t1.y);
      // (i h) t2 = (-h t2.y, h t2.x)
      a[0] = b0;
      a[1] = make_float2(m.x - h * t2.y, m.y + h * t2.x);
      a[2] = make_float2(m.x + h * t2.y, m.y - h * t2.x);
    } else {
      // radix 4: W_4 = exp(s 2 pi i / 4) = s*i (s = -1: -i)
      const float2 t0 = make_float2(a[0].x + a[2].x, a[0].y + a[2].y);
      const float2 t1 = make_float2(a[0].x - a[2].x, a[0].y - a[2].y);
      const float2 t2 = make_float2(a[1].x + a[3].x, a[1].y + a[3].y);
      const float2 t3 = make_float2(a[1].x - a[3].x, a[1].y - a[3].y);
      // t3 * (s i) = (-s t3.y, s t3.x)
      const float2 u3 = make_float2(-s * t3.y, s * t3.x);
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

// in-place N-point DFT (N = 16 or 32), natural order in and out
template <int N>
__device__ __forceinline__ void reg_fft(float2 (&v)[N], float s) {
  float2 t[N];
  if (N == 16) {
    reg_stage<N, 1, 4>(v, t, s);
    reg_stage<N, 4, 4>(t, v, s);
  } else {
    reg_stage<N, 1, 4>(v, t, s);
    reg_stage<N, 4, 4>(t, v, s);
    reg_stage<N, 16, 2>(v, t, s);
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = t[i];
  }
}

__device__ __forceinline__ float deapod_r(const FftArgs& a, int i0, int64_t rest) {
  float f = a.dinv0[i0];
  const int i1 = (int)(rest % a.N1);
  f *= a.dinv1[i1];
  if (a.d == 3) f *= a.dinv2[(int)(rest / a.N1)];
  return f;
}

__device__ __forceinline__ bool inside_mask_r(const FftArgs& a, int i0, int64_t rest) {
  const int nh = a.mask_nh;
  const int i1 = (int)(rest % a.N1);
  int64_t s = (int64_t)(i0 - nh) * (i0 - nh) + (int64_t)(i1 - nh) * (i1 - nh);
  if (a.d == 3) {
    const int i2 = (int)(rest / a.N1);
    s += (int64_t)(i2 - nh) * (i2 - nh);
  }
  return s <= (int64_t)nh * nh;
}

constexpr int kRegLines = 16;

// sum over the 256-thread block (every thread calls it)
__device__ __forceinline__ double block_sum_d(double v, double* red) {
  v = wave_sum_d(v);
  lds_barrier();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  lds_barrier();
  return red[0] + red[1] + red[2] + red[3];
}

// exchange layout: element (row k1, column j) of line q at
// ex[k1 * XS + j * 16 + q], XS = 16*16 + 8 (pad), so a wave's 64 lanes
// (16 lines x 4 columns) touch 64 consecutive slots on writes
template <int N1>
__device__ __forceinline__ int xidx(int k1, int j, int q) {
  return k1 * (16 * 16 + 8) + j * 16 + q;
}

}  // namespace

// XR: rows of the exchange buffer.  XR = N1 exchanges everything at once;
// XR = 16 exchanges in N1/16 rounds of 16 rows (half the LDS at n = 512, so
// more blocks fit on a CU, at the price of extra barriers)
template <int N1, int LD, int ST, int CONV, int XR>
__global__ void __launch_bounds__(256) k_fft_reg(FftArgs a) {
  constexpr int n = 16 * N1;
  constexpr int RPT = N1 / 16;   // 16-point rows per thread after the exchange
  static_assert(XR == N1 || XR == 16, "exchange rows");
  __shared__ float2 ex[XR * (16 * 16 + 8)];
  __shared__ float2 tw[n];
  __shared__ double red[4];
  const int tid = threadIdx.x;
  const int q = tid & 15, j = tid >> 4;
  const int64_t nib = (a.inner + kRegLines - 1) / kRegLines;
  const int64_t o = blockIdx.x / nib;
  const int64_t q0 = (blockIdx.x - o * nib) * kRegLines;
  const int64_t f = o / a.outer_per_frame;
  if (a.fdone && a.fdone[f]) return;
  const int64_t rest = o - f * a.outer_per_frame;
  const int nq = (int)min((int64_t)kRegLines, a.inner - q0);
  const bool live = q < nq;
  const int64_t qq = q0 + q;
  const float s = (float)a.sign;

  for (int i = tid; i < n; i += 256) tw[i] = a.tw[i];

  // ---- load positions pos = j + 16 k (image index p with the zero pad) ----
  float2 v[N1];
  float ppacc = 0.f;   // ||p_new||^2 over this row (coil-0 lanes of the fused CG update)
  const bool pad = a.pad_in != 0;
#pragma unroll
  for (int k = 0; k < N1; ++k) {
    const int pos = j + 16 * k;
    int p = pos;
    if (pad) {
      p = pos + a.nin / 2;
      if (p >= n) p -= n;
    }
    float2 t = make_float2(0.f, 0.f);
    if (live && p < a.nin) {
      if (LD == 1 || LD == 4) {
        const int64_t i = p + (int64_t)a.N0 * rest;
        float2 xv;
        if (a.pold) {   // p_new = r + beta p_old (CGLS.hpp:55-56), fused
          const float be = (float)a.cgst[f].beta;
          const float2 rr = a.r[f * a.Ntot + i], po = a.pold[f * a.Ntot + i];
          xv = make_float2(rr.x + be * po.x, rr.y + be * po.y);
          if (qq == 0) {
            a.pnew[f * a.Ntot + i] = xv;
            ppacc += xv.x * xv.x + xv.y * xv.y;
          }
        } else {
          xv = a.x[f * a.Ntot + i];
        }
        t = cmul2(xv, a.S[f * a.Sfs + i * a.C + qq]);
        if (LD == 1) {
          const float dv = deapod_r(a, p, rest);
          t = make_float2(t.x * dv, t.y * dv);
        }
      } else {
        t = a.in[(o * a.nin + p) * a.inner + qq];
        if (LD >= 2) {
          float dv = deapod_r(a, p, rest);
          if (LD == 3 && !inside_mask_r(a, p, rest)) dv = 0.f;
          t = make_float2(t.x * dv, t.y * dv);
        }
      }
    }
    v[k] = t;
  }
  if ((LD == 1 || LD == 4) && a.pp_part && q0 == 0) {   // block-uniform condition
    const double t = block_sum_d(ppacc, red);
    if (tid == 0) a.pp_part[f * a.outer_per_frame + rest] = t;
  }
  lds_barrier();   // twiddle table

  // ---- N1-point DFT over k, twiddle W_n^(s j k1), exchange ----
  reg_fft<N1>(v, s);
#pragma unroll
  for (int k1 = 1; k1 < N1; ++k1) {
    const float2 w = tw[j * k1];
    v[k1] = cmul2(v[k1], make_float2(w.x, s < 0 ? w.y : -w.y));
  }
  // rows k1 = j + 16 r: 16-point DFT over the column index
  float2 X[RPT][16];
  if (XR == N1) {
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) ex[xidx<N1>(k1, j, q)] = v[k1];
    lds_barrier();
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
#pragma unroll
      for (int jj = 0; jj < 16; ++jj) X[r][jj] = ex[xidx<N1>(j + 16 * r, jj, q)];
      reg_fft<16>(X[r], s);
    }
  } else {
#pragma unroll
    for (int r = 0; r < RPT; ++r) {   // round r holds rows 16 r .. 16 r + 15
      if (r) lds_barrier();
#pragma unroll
      for (int kk = 0; kk < 16; ++kk) ex[xidx<N1>(kk, j, q)] = v[16 * r + kk];
      lds_barrier();
#pragma unroll
      for (int jj = 0; jj < 16; ++jj) X[r][jj] = ex[xidx<N1>(j, jj, q)];
      reg_fft<16>(X[r], s);
    }
  }
  // thread holds bins P = k1 + N1 k2, k1 = j + 16 r, k2 = 0..15

  if (CONV) {
    // multiply by the real kernel spectrum, then the inverse transform
    const float* th = a.That + (f * a.conv_cells + qq / a.C) * n;
#pragma unroll
    for (int r = 0; r < RPT; ++r)
#pragma unroll
      for (int k2 = 0; k2 < 16; ++k2) {
        const float t = live ? th[(j + 16 * r) + N1 * k2] : 0.f;
        X[r][k2] = make_float2(X[r][k2].x * t, X[r][k2].y * t);
      }
    const float si = -s;
#pragma unroll
    for (int r = 0; r < RPT; ++r) reg_fft<16>(X[r], si);   // over k2 -> index j'
    if (XR == N1) {
      lds_barrier();   // every thread has read its rows of ex
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        const int k1 = j + 16 * r;
#pragma unroll
        for (int jj = 0; jj < 16; ++jj) {
          const float2 w = tw[jj * k1];
          ex[xidx<N1>(k1, jj, q)] = cmul2(X[r][jj], make_float2(w.x, si < 0 ? w.y : -w.y));
        }
      }
      lds_barrier();
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) v[k1] = ex[xidx<N1>(k1, j, q)];
    } else {
#pragma unroll
      for (int r = 0; r < RPT; ++r) {
        const int k1 = j + 16 * r;
        lds_barrier();   // the previous round's reads are done
#pragma unroll
        for (int jj = 0; jj < 16; ++jj) {
          const float2 w = tw[jj * k1];
          ex[xidx<N1>(j, jj, q)] = cmul2(X[r][jj], make_float2(w.x, si < 0 ? w.y : -w.y));
        }
        lds_barrier();
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) v[16 * r + kk] = ex[xidx<N1>(kk, j, q)];
      }
    }
    reg_fft<N1>(v, si);   // over k1 -> k; point j + 16 k
    // ---- store positions j + 16 k (crop) ----
    const bool crop = a.crop_out != 0;
#pragma unroll
    for (int k = 0; k < N1; ++k) {
      const int pos = j + 16 * k;
      int p = pos;
      if (crop) {
        p = pos + a.nout / 2;
        if (p >= n) p -= n;
      }
      if (live && p < a.nout) a.out[(o * a.nout + p) * a.inner + qq] = v[k];
    }
    return;
  }

  // ---- store bins P (crop: output index p = P + nout/2 mod n) ----
  const bool crop = a.crop_out != 0;
  if (ST == 2) {
    const int64_t g = q0 / kRegLines;
    float2* outp = a.out + (f * nib + g) * a.Ntot + (int64_t)a.N0 * rest;
    const float2* pd = a.pdot ? a.pdot + f * a.Ntot + (int64_t)a.N0 * rest : nullptr;
    double pqacc = 0.0;
#pragma unroll
    for (int r = 0; r < RPT; ++r)
#pragma unroll
      for (int k2 = 0; k2 < 16; ++k2) {
        const int P = (j + 16 * r) + N1 * k2;
        int p = P + a.nout / 2;
        if (p >= n) p -= n;
        float2 c = make_float2(0.f, 0.f);
        if (live && p < a.nout) {
          const float2 sv = a.S[f * a.Sfs + (p + (int64_t)a.N0 * rest) * a.C + qq];
          const float2 xv = X[r][k2];
          c = make_float2(sv.x * xv.x + sv.y * xv.y, sv.x * xv.y - sv.y * xv.x);
        }
        c = make_float2(row16_sum(c.x), row16_sum(c.y));   // over the 16 coils (lanes q of a row)
        if (q == 0 && p < a.nout) {
          outp[p] = c;
          if (pd) {
            const float2 pv = pd[p];
            pqacc += (double)pv.x * c.x + (double)pv.y * c.y;
          }
        }
      }
    if (a.pq_part) {   // block-uniform
      const double t = block_sum_d(pqacc, red);
      if (tid == 0) a.pq_part[(f * nib + g) * a.outer_per_frame + rest] = t;
    }
    return;
  }
#pragma unroll
  for (int r = 0; r < RPT; ++r)
#pragma unroll
    for (int k2 = 0; k2 < 16; ++k2) {
      const int P = (j + 16 * r) + N1 * k2;
      int p = P;
      if (crop) {
        p = P + a.nout / 2;
        if (p >= n) p -= n;
      }
      if (live && p < a.nout) {
        float2 val = X[r][k2];
        if (ST == 1) {
          const float dv = deapod_r(a, p, rest);
          val = make_float2(val.x * dv, val.y * dv);
        }
        a.out[(o * a.nout + p) * a.inner + qq] = val;
      }
    }
}

// ----------------------------------------------------------------------------
// 512-point lines as 8 x 8 x 8: a 1024-thread block takes 16 lines, 64 threads
// per line, 8 points per thread (p = t + 64 k).  Three in-register 8-point
// DFTs with two LDS exchanges between them (four-step applied twice:
// 512 = 8 x 64, 64 = 8 x 8).  Twice the LDS traffic of the 16-thread layout,
// but 8 points per thread keep VGPRs low, so a CU runs 32 waves instead of
// 8-12 and every pass has four times as many threads in flight.
namespace {

constexpr int kR8Row = 144;   // floats2 per exchange row: 8 columns x 16 lines + pad

// 8-point DFT in registers, natural order (radix 2 x 4 Stockham)
__device__ __forceinline__ void dft8(float2 (&v)[8], float s) {
  float2 t[8];
  reg_stage<8, 1, 2>(v, t, s);
  reg_stage<8, 2, 4>(t, v, s);
}

__device__ __forceinline__ float2 twid(const float2* tw, int idx, float s) {
  const float2 w = tw[idx];
  return make_float2(w.x, s < 0 ? w.y : -w.y);
}

__device__ __forceinline__ double block_sum_1024(double v, double* red) {
  v = wave_sum_d(v);
  lds_barrier();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  lds_barrier();
  double t = 0.0;
#pragma unroll
  for (int w = 0; w < 16; ++w) t += red[w];
  return t;
}

}  // namespace

// small DFT of N1 = 1..8 points in registers (natural order)
template <int N1>
__device__ __forceinline__ void dftN(float2 (&v)[N1], float s) {
  if constexpr (N1 == 2) {
    float2 t[2];
    reg_stage<2, 1, 2>(v, t, s);
    v[0] = t[0]; v[1] = t[1];
  } else if constexpr (N1 == 3) {
    float2 t[3];
    reg_stage<3, 1, 3>(v, t, s);
    v[0] = t[0]; v[1] = t[1]; v[2] = t[2];
  } else if constexpr (N1 == 4) {
    float2 t[4];
    reg_stage<4, 1, 4>(v, t, s);
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = t[i];
  } else if constexpr (N1 == 6) {
    float2 t[6];
    reg_stage<6, 1, 2>(v, t, s);
    reg_stage<6, 2, 3>(t, v, s);
  } else if constexpr (N1 == 8) {
    dft8(v, s);
  }
}

// n = 64 N1 (N1 in {2, 3, 4, 6, 8}): 64 threads per line hold N1 points each
// (p = t + 64 k) for an N1-point DFT, then the 64-point DFT over t runs as
// 8 x 8 on the first 8 N1 threads of the line (8 points each) with two
// exchanges.  Same fusions as the other FFT kernels.
template <int N1, int LD, int ST, int CONV>
__global__ void __launch_bounds__(1024) k_fft_r8(FftArgs a) {
  constexpr int n = 64 * N1;
  constexpr int ACT = 8 * N1;   // threads per line busy in the 8 x 8 stages
  __shared__ float2 ex[ACT * kR8Row];
  __shared__ float2 tw[n];
  __shared__ double red[16];
  const int tid = threadIdx.x;
  const int q = tid & 15, t = tid >> 4;   // line in the block, thread in the line
  const int lo = t & 7, hi = t >> 3;
  const int64_t nib = (a.inner + kRegLines - 1) / kRegLines;
  const int64_t o = blockIdx.x / nib;
  const int64_t q0 = (blockIdx.x - o * nib) * kRegLines;
  const int64_t f = o / a.outer_per_frame;
  if (a.fdone && a.fdone[f]) return;
  const int64_t rest = o - f * a.outer_per_frame;
  const int nq = (int)min((int64_t)kRegLines, a.inner - q0);
  const bool live = q < nq;
  const int64_t qq = q0 + q;
  const float s = (float)a.sign;

  for (int i = tid; i < n; i += 1024) tw[i] = a.tw[i];

  // ---- load positions t + 64 k ----
  float2 u[N1];
  float ppacc = 0.f;
  const bool pad = a.pad_in != 0;
#pragma unroll
  for (int k = 0; k < N1; ++k) {
    int p = t + 64 * k;
    if (pad) {
      p += a.nin / 2;
      if (p >= n) p -= n;
    }
    float2 x = make_float2(0.f, 0.f);
    if (live && p < a.nin) {
      if (LD == 1 || LD == 4) {
        const int64_t i = p + (int64_t)a.N0 * rest;
        float2 xv;
        if (a.pold) {   // p_new = r + beta p_old (CGLS.hpp:55-56), fused
          const float be = (float)a.cgst[f].beta;
          const float2 rr = a.r[f * a.Ntot + i], po = a.pold[f * a.Ntot + i];
          xv = make_float2(rr.x + be * po.x, rr.y + be * po.y);
          if (qq == 0) {
            a.pnew[f * a.Ntot + i] = xv;
            ppacc += xv.x * xv.x + xv.y * xv.y;
          }
        } else {
          xv = a.x[f * a.Ntot + i];
        }
        x = cmul2(xv, a.S[f * a.Sfs + i * a.C + qq]);
        if (LD == 1) {
          const float dv = deapod_r(a, p, rest);
          x = make_float2(x.x * dv, x.y * dv);
        }
      } else {
        x = a.in[(o * a.nin + p) * a.inner + qq];
        if (LD >= 2) {
          float dv = deapod_r(a, p, rest);
          if (LD == 3 && !inside_mask_r(a, p, rest)) dv = 0.f;
          x = make_float2(x.x * dv, x.y * dv);
        }
      }
    }
    u[k] = x;
  }
  if ((LD == 1 || LD == 4) && a.pp_part && q0 == 0) {   // block-uniform
    const double r = block_sum_1024(ppacc, red);
    if (tid == 0) a.pp_part[f * a.outer_per_frame + rest] = r;
  }
  lds_barrier();   // twiddle table

  // stage 1: DFT over k (t = lo + 8 hi), twiddle W_n^(t k1)
  if (N1 > 1) dftN<N1>(u, s);
#pragma unroll
  for (int k1 = 1; k1 < N1; ++k1) u[k1] = cmul2(u[k1], twid(tw, t * k1, s));
  // exchange 1: element (k1, lo, hi) -> row 8 k1 + lo, column hi
#pragma unroll
  for (int k1 = 0; k1 < N1; ++k1) ex[(8 * k1 + lo) * kR8Row + hi * 16 + q] = u[k1];
  lds_barrier();
  const bool act = t < ACT;   // thread t = 8 k1 + j1 takes row t: values over j2
  float2 v[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) v[c] = act ? ex[t * kR8Row + c * 16 + q] : make_float2(0.f, 0.f);
  // stage 2: DFT over j2 -> m1, twiddle W_64^(j1 m1) = W_n^(N1 j1 m1)
  dft8(v, s);
#pragma unroll
  for (int m1 = 1; m1 < 8; ++m1) v[m1] = cmul2(v[m1], twid(tw, N1 * lo * m1, s));
  lds_barrier();
  // exchange 2: element (k1 = hi, m1, j1 = lo) -> row 8 k1 + m1, column j1
  if (act) {
#pragma unroll
    for (int m1 = 0; m1 < 8; ++m1) ex[(8 * hi + m1) * kR8Row + lo * 16 + q] = v[m1];
  }
  lds_barrier();
#pragma unroll
  for (int c = 0; c < 8; ++c) v[c] = act ? ex[t * kR8Row + c * 16 + q] : make_float2(0.f, 0.f);
  // stage 3: DFT over j1 -> m2; thread t = 8 k1 + m1 holds bins k1 + N1 m1 + 8 N1 m2
  dft8(v, s);
  const int binbase = hi + N1 * lo;

  if (CONV) {
    const float* th = a.That + (f * a.conv_cells + qq / a.C) * n;
#pragma unroll
    for (int m2 = 0; m2 < 8; ++m2) {
      const float w = (live && act) ? th[binbase + 8 * N1 * m2] : 0.f;
      v[m2] = make_float2(v[m2].x * w, v[m2].y * w);
    }
    const float si = -s;
    // the inverse runs the stages backwards with conjugate twiddles
    dft8(v, si);                                            // over m2 -> j1
    lds_barrier();
    if (act) {
#pragma unroll
      for (int j1 = 0; j1 < 8; ++j1) ex[t * kR8Row + j1 * 16 + q] = v[j1];
    }
    lds_barrier();
    // thread t = 8 k1 + j1 gathers over m1
#pragma unroll
    for (int m1 = 0; m1 < 8; ++m1) v[m1] = act ? ex[(8 * hi + m1) * kR8Row + lo * 16 + q] : make_float2(0.f, 0.f);
#pragma unroll
    for (int m1 = 1; m1 < 8; ++m1) v[m1] = cmul2(v[m1], twid(tw, N1 * lo * m1, si));
    dft8(v, si);                                            // over m1 -> j2
    lds_barrier();
    if (act) {
#pragma unroll
      for (int j2 = 0; j2 < 8; ++j2) ex[t * kR8Row + j2 * 16 + q] = v[j2];
    }
    lds_barrier();
    // thread t = j1 + 8 j2 (lo = j1, hi = j2) gathers over k1
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) u[k1] = ex[(8 * k1 + lo) * kR8Row + hi * 16 + q];
#pragma unroll
    for (int k1 = 1; k1 < N1; ++k1) u[k1] = cmul2(u[k1], twid(tw, t * k1, si));
    if (N1 > 1) dftN<N1>(u, si);                            // over k1 -> k: point t + 64 k
    const bool crop = a.crop_out != 0;
#pragma unroll
    for (int k = 0; k < N1; ++k) {
      int p = t + 64 * k;
      if (crop) {
        p += a.nout / 2;
        if (p >= n) p -= n;
      }
      if (live && p < a.nout) a.out[(o * a.nout + p) * a.inner + qq] = u[k];
    }
    return;
  }

  const bool crop = a.crop_out != 0;
  if (ST == 2) {
    const int64_t g = q0 / kRegLines;
    float2* outp = a.out + (f * nib + g) * a.Ntot + (int64_t)a.N0 * rest;
    const float2* pd = a.pdot ? a.pdot + f * a.Ntot + (int64_t)a.N0 * rest : nullptr;
    double pqacc = 0.0;
#pragma unroll
    for (int m2 = 0; m2 < 8; ++m2) {
      int p = binbase + 8 * N1 * m2 + a.nout / 2;
      if (p >= n) p -= n;
      float2 c = make_float2(0.f, 0.f);
      if (live && act && p < a.nout) {
        const float2 sv = a.S[f * a.Sfs + (p + (int64_t)a.N0 * rest) * a.C + qq];
        c = make_float2(sv.x * v[m2].x + sv.y * v[m2].y, sv.x * v[m2].y - sv.y * v[m2].x);
      }
      c = make_float2(row16_sum(c.x), row16_sum(c.y));   // over the 16 coils (lanes q of a row)
      if (q == 0 && act && p < a.nout) {
        outp[p] = c;
        if (pd) {
          const float2 pv = pd[p];
          pqacc += (double)pv.x * c.x + (double)pv.y * c.y;
        }
      }
    }
    if (a.pq_part) {   // block-uniform
      const double r = block_sum_1024(pqacc, red);
      if (tid == 0) a.pq_part[(f * nib + g) * a.outer_per_frame + rest] = r;
    }
    return;
  }
#pragma unroll
  for (int m2 = 0; m2 < 8; ++m2) {
    int p = binbase + 8 * N1 * m2;
    if (crop) {
      p += a.nout / 2;
      if (p >= n) p -= n;
    }
    if (live && act && p < a.nout) {
      float2 val = v[m2];
      if (ST == 1) {
        const float dv = deapod_r(a, p, rest);
        val = make_float2(val.x * dv, val.y * dv);
      }
      a.out[(o * a.nout + p) * a.inner + qq] = val;
    }
  }
}

// lines of 64 N1 points on the 64-threads-per-line kernel unless CGS_FFT_R8=0
static bool fft_r8_enabled() {
  static const bool on = [] {
    const char* e = std::getenv("CGS_FFT_R8");
    return !(e && std::atoi(e) == 0);
  }();
  return on;
}

bool fft_reg_supported(const FftArgs& a, int ld_mode, int st_mode) {
  const bool r8 = fft_r8_enabled() && (a.n == 128 || a.n == 192 || a.n == 384);
  if (!(a.n == 256 || a.n == 512 || r8)) return false;
  if (a.inner < kRegLines) return false;
  if (st_mode == 2 && a.That) return false;
  (void)ld_mode;
  return true;
}

// exchange in rounds of 16 rows at n = 512 unless CGS_FFT_XFULL=1
static bool fft_reg_full_exchange() {
  static const bool full = [] {
    const char* e = std::getenv("CGS_FFT_XFULL");
    return e && std::atoi(e) == 1;
  }();
  return full;
}

void launch_fft_reg(const FftArgs& a, int ld_mode, int st_mode, hipStream_t s) {
  const int64_t nib = (a.inner + kRegLines - 1) / kRegLines;
  const dim3 grid((unsigned)(nib * a.outers));
  const int conv = a.That != nullptr;
  const bool full = fft_reg_full_exchange();
  // the Toeplitz convolution pass (forward, multiply, inverse) stays on the
  // 16-threads-per-line kernel at n = 256 / 512, where it measured faster
  const bool r8 = fft_r8_enabled() && (a.n == 512 || a.n == 384 || a.n == 256 || a.n == 192 || a.n == 128) &&
                  !(conv && (a.n == 256 || a.n == 512));
  if (r8) {
#define CGS_FFT8_CASE(L, S_, CV)                                                        \
    if (ld_mode == L && st_mode == S_ && conv == CV) {                                  \
      switch (a.n) {                                                                    \
        case 512: CGS_LAUNCH((k_fft_r8<8, L, S_, CV>), grid, dim3(1024), 0, s, a); break; \
        case 384: CGS_LAUNCH((k_fft_r8<6, L, S_, CV>), grid, dim3(1024), 0, s, a); break; \
        case 256: CGS_LAUNCH((k_fft_r8<4, L, S_, CV>), grid, dim3(1024), 0, s, a); break; \
        case 192: CGS_LAUNCH((k_fft_r8<3, L, S_, CV>), grid, dim3(1024), 0, s, a); break; \
        default:  CGS_LAUNCH((k_fft_r8<2, L, S_, CV>), grid, dim3(1024), 0, s, a); break; \
      }                                                                                 \
      return;                                                                           \
    }
    CGS_FFT8_CASE(0, 0, 0)
    CGS_FFT8_CASE(1, 0, 0)
    CGS_FFT8_CASE(2, 0, 0)
    CGS_FFT8_CASE(3, 0, 0)
    CGS_FFT8_CASE(4, 0, 0)
    CGS_FFT8_CASE(0, 1, 0)
    CGS_FFT8_CASE(0, 0, 1)
    CGS_FFT8_CASE(0, 2, 0)
#undef CGS_FFT8_CASE
  }
#define CGS_FFTR_CASE(L, S_, CV)                                                      \
  if (ld_mode == L && st_mode == S_ && conv == CV) {                                  \
    if (a.n == 512 && full)                                                           \
      CGS_LAUNCH((k_fft_reg<32, L, S_, CV, 32>), grid, dim3(256), 0, s, a);  \
    else if (a.n == 512)                                                              \
      CGS_LAUNCH((k_fft_reg<32, L, S_, CV, 16>), grid, dim3(256), 0, s, a);  \
    else                                                                              \
      CGS_LAUNCH((k_fft_reg<16, L, S_, CV, 16>), grid, dim3(256), 0, s, a);  \
    return;                                                                           \
  }
  CGS_FFTR_CASE(0, 0, 0)
  CGS_FFTR_CASE(1, 0, 0)
  CGS_FFTR_CASE(2, 0, 0)
  CGS_FFTR_CASE(3, 0, 0)
  CGS_FFTR_CASE(4, 0, 0)
  CGS_FFTR_CASE(0, 1, 0)
  CGS_FFTR_CASE(0, 0, 1)
  CGS_FFTR_CASE(0, 2, 0)
#undef CGS_FFTR_CASE
}

}  // namespace cgs
