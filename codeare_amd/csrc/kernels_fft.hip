// kernels_fft.hip -- batched 1D FFT passes over one axis of the oversampled grid.
//
// One pass transforms every line along one axis of an array viewed as
// [outer][P][inner] (inner contiguous).  A block stages `lb` adjacent lines in
// LDS ([lb][n + n/16 + 1]), so global loads and stores move lb*8 contiguous bytes per
// position; each line is then transformed by one wave with a mixed-radix
// Stockham autosort (radix 2,3,4,8) in place in LDS.  Global loads are issued
// eight at a time before any of them is waited on, and the twiddle table is
// staged in LDS, so a block pays about one memory latency per phase.
//
// Forward passes read image-index order (i in [0, nin), frequency i - nin/2,
// zero-padded into bin (i - nin/2) mod n) and write grid order; adjoint passes
// read grid order and write image-index order (crop).  Pass 1 of the forward
// transform fuses the deapodisation 1/phi_hat, the zero padding and either the
// coil multiply S_c .* x (E) or the damping mask w_hat (the CGNR's A(w_hat p));
// the last adjoint pass fuses the crop and the deapodisation.  The Toeplitz
// pass runs forward FFT, spectrum multiply and inverse FFT on the same lines.
#include "cgs_internal.hpp"

namespace cgs {

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// sign * i * z
__device__ __forceinline__ float2 cmuli(float2 z, int sign) {
  return sign > 0 ? make_float2(-z.y, z.x) : make_float2(z.y, -z.x);
}

template <int R>
__device__ __forceinline__ void dft(float2* v, int sign);

template <>
__device__ __forceinline__ void dft<2>(float2* v, int) {
  float2 a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <>
__device__ __forceinline__ void dft<3>(float2* v, int sign) {
  const float c = -0.5f, s = (sign > 0 ? 1.0f : -1.0f) * 0.86602540378443864676f;
  float2 a = v[0], b = v[1], d = v[2];
  float2 sum = cadd(b, d), dif = csub(b, d);
  float2 m = make_float2(a.x + c * sum.x, a.y + c * sum.y);
  float2 t = make_float2(-s * dif.y, s * dif.x);   // i*s*dif
  v[0] = cadd(a, sum);
  v[1] = cadd(m, t);
  v[2] = csub(m, t);
}

template <>
__device__ __forceinline__ void dft<4>(float2* v, int sign) {
  float2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
  float2 b0 = cadd(v[1], v[3]), b1 = cmuli(csub(v[1], v[3]), sign);
  v[0] = cadd(a0, b0);
  v[2] = csub(a0, b0);
  v[1] = cadd(a1, b1);
  v[3] = csub(a1, b1);
}

template <>
__device__ __forceinline__ void dft<8>(float2* v, int sign) {
  float2 e[4] = {v[0], v[2], v[4], v[6]};
  float2 o[4] = {v[1], v[3], v[5], v[7]};
  dft<4>(e, sign);
  dft<4>(o, sign);
  const float h = 0.70710678118654752440f;
  const float sg = sign > 0 ? 1.0f : -1.0f;
  // W8^1 = (1 + sg i)/sqrt2, W8^2 = sg i, W8^3 = (-1 + sg i)/sqrt2
  float2 o1 = make_float2(h * (o[1].x - sg * o[1].y), h * (o[1].y + sg * o[1].x));
  float2 o2 = cmuli(o[2], sign);
  float2 o3 = make_float2(h * (-o[3].x - sg * o[3].y), h * (-o[3].y + sg * o[3].x));
  v[0] = cadd(e[0], o[0]);
  v[4] = csub(e[0], o[0]);
  v[1] = cadd(e[1], o1);
  v[5] = csub(e[1], o1);
  v[2] = cadd(e[2], o2);
  v[6] = csub(e[2], o2);
  v[3] = cadd(e[3], o3);
  v[7] = csub(e[3], o3);
}

// One Stockham stage of radix R on one line held in LDS, done by one wave.
// Lane j handles butterflies j, j+64, ...; all of its reads are issued before
// its writes, and a wave's LDS operations execute in order, so the stage runs
// in place without a block barrier.  MB = butterflies per lane at most.
// j / Ns is formed as umulhi(j, ceil(2^32 / Ns)) (exact for j < 2^16).
// Element L of a line sits at pidx(L) = L + L/16: the pad slot every 16
// complex values spreads the stride-R writes of the early stages over banks.
__device__ __forceinline__ int pidx(int L) { return L + (L >> 4); }

template <int R, int MB>
__device__ __forceinline__ void wave_stage(float2* line, int n, int Ns, uint32_t magic, int lane, int sign,
                                           const float2* __restrict__ tw) {
  const int nb = n / R;
  const int tstep0 = n / (Ns * R);
  float2 v[MB][R];
#pragma unroll
  for (int k = 0; k < MB; ++k) {
    const int j = lane + 64 * k;
    if (j < nb) {
      const int kk = (Ns == 1) ? 0 : j - (int)__umulhi((uint32_t)j, magic) * Ns;
      const int tstep = tstep0 * kk;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        float2 x = line[pidx(j + r * nb)];
        if (r > 0) {
          float2 w = tw[r * tstep];
          if (sign > 0) w.y = -w.y;
          x = cmul(x, w);
        }
        v[k][r] = x;
      }
      dft<R>(v[k], sign);
    }
  }
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int k = 0; k < MB; ++k) {
    const int j = lane + 64 * k;
    if (j < nb) {
      const int jq = (Ns == 1) ? j : (int)__umulhi((uint32_t)j, magic);
      const int kk = j - jq * Ns;
      const int d0 = jq * Ns * R + kk;
#pragma unroll
      for (int r = 0; r < R; ++r) line[pidx(d0 + r * Ns)] = v[k][r];
    }
  }
  __builtin_amdgcn_wave_barrier();
}

template <int MAXN>
__device__ __forceinline__ void line_fft(float2* line, const FftArgs& a, int lane, int sign, const float2* tw) {
  int Ns = 1;
  for (int s = 0; s < a.nst; ++s) {
    const int R = a.radix[s];
    const uint32_t mg = a.ns_magic[s];
    if (R == 8) wave_stage<8, (MAXN / 8 + 63) / 64>(line, a.n, Ns, mg, lane, sign, tw);
    else if (R == 4) wave_stage<4, (MAXN / 4 + 63) / 64>(line, a.n, Ns, mg, lane, sign, tw);
    else if (R == 2) wave_stage<2, (MAXN / 2 + 63) / 64>(line, a.n, Ns, mg, lane, sign, tw);
    else wave_stage<3, (MAXN / 3 + 63) / 64>(line, a.n, Ns, mg, lane, sign, tw);
    Ns *= R;
  }
}

__device__ __forceinline__ float deapod(const FftArgs& a, int i0, int64_t rest) {
  float f = a.dinv0[i0];
  const int i1 = (int)(rest % a.N1);
  f *= a.dinv1[i1];
  if (a.d == 3) f *= a.dinv2[(int)(rest / a.N1)];
  return f;
}

__device__ __forceinline__ bool inside_mask(const FftArgs& a, int i0, int64_t rest) {
  const int nh = a.mask_nh;
  const int i1 = (int)(rest % a.N1);
  int64_t s = (int64_t)(i0 - nh) * (i0 - nh) + (int64_t)(i1 - nh) * (i1 - nh);
  if (a.d == 3) {
    const int i2 = (int)(rest / a.N1);
    s += (int64_t)(i2 - nh) * (i2 - nh);
  }
  return s <= (int64_t)nh * nh;
}

constexpr int kUnroll = 8;   // global loads in flight per thread in the staging loops

// LD: 0 plain, 1 E (x .* S_c .* D), 2 plain .* D, 3 plain .* D .* w_hat, 4 x .* S_c
// ST: 0 plain, 1 .* D, 2 crop + coil combine: the block's lb lines are lb coils
//     of one image row (axis-0 pass, inner = C); it writes
//     sum_{c in block} conj(S_c) X_c to partial image g = q0 / lb:
//     out[(f*G + g)*Ntot + i], G = ceil(C / lb)
// CONV: forward FFT, multiply by the Toeplitz kernel spectrum, inverse FFT
template <int LD, int ST, int MAXN, int CONV>
__global__ void __launch_bounds__(kFftThreads) k_fft(FftArgs a) {
  extern __shared__ float2 lds[];
  const int tid = threadIdx.x;
  const int lbl = a.lb_log2, lb = 1 << lbl, n = a.n;
  const int64_t nib = (a.inner + lb - 1) >> lbl;
  const int64_t o = blockIdx.x / nib;
  const int64_t q0 = (blockIdx.x - o * nib) << lbl;
  const int64_t f = o / a.outer_per_frame;
  if (a.fdone && a.fdone[f]) return;
  const int64_t rest = o - f * a.outer_per_frame;
  const int nq = (int)min((int64_t)lb, a.inner - q0);
  float2* tw = lds + (size_t)lb * a.ldn;

  for (int i = tid; i < n; i += kFftThreads) tw[i] = a.tw[i];

  // ---- load (with zero padding for forward passes) ----
  const bool pad = a.pad_in != 0;   // forward passes read image-index order
  const int total = lb * n;
  for (int b0 = tid; b0 < total; b0 += kFftThreads * kUnroll) {
    float2 v[kUnroll];
    int dst[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int idx = b0 + u * kFftThreads;
      v[u] = make_float2(0.f, 0.f);
      dst[u] = -1;
      if (idx < total) {
        const int q = idx & (lb - 1), pos = idx >> lbl;
        dst[u] = q * a.ldn + pidx(pos);
        int p = pos;
        if (pad) {   // bin pos holds image index p = pos + nin/2 (mod n) when p < nin
          p = pos + a.nin / 2;
          if (p >= n) p -= n;
        }
        if (q < nq && p < a.nin) {
          const int64_t qq = q0 + q;
          if (LD == 1 || LD == 4) {
            const int64_t i = p + (int64_t)a.N0 * rest;
            float2 xv;
            if (a.pold) {   // p_new = r + beta p_old (CGLS.hpp:55-56), fused
              const float be = (float)a.cgst[f].beta;
              const float2 rr = a.r[f * a.Ntot + i], po = a.pold[f * a.Ntot + i];
              xv = make_float2(rr.x + be * po.x, rr.y + be * po.y);
              if (qq == 0) a.pnew[f * a.Ntot + i] = xv;
            } else {
              xv = a.x[f * a.Ntot + i];
            }
            const float2 sv = a.S[f * a.Sfs + i * a.C + qq];
            float2 t = cmul(xv, sv);
            if (LD == 1) {
              const float dv = deapod(a, p, rest);
              t = make_float2(t.x * dv, t.y * dv);
            }
            v[u] = t;
          } else {
            float2 t = a.in[(o * a.nin + p) * a.inner + qq];
            if (LD >= 2) {
              float dv = deapod(a, p, rest);
              if (LD == 3 && !inside_mask(a, p, rest)) dv = 0.f;
              t = make_float2(t.x * dv, t.y * dv);
            }
            v[u] = t;
          }
        }
      }
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u)
      if (dst[u] >= 0) lds[dst[u]] = v[u];
  }
  lds_barrier();

  // ---- Stockham stages: wave w transforms lines w, w+4, ... ----
  {
    const int lane = tid & 63, wave = tid >> 6;
    for (int q = wave; q < nq; q += kFftThreads / 64) {
      float2* line = lds + q * a.ldn;
      if (CONV) {
        line_fft<MAXN>(line, a, lane, -1, tw);
        const float* th = a.That + (f * a.conv_cells + (q0 + q) / a.C) * n;
        for (int pos = lane; pos < n; pos += 64) {
          const float t = th[pos];
          float2 v = line[pidx(pos)];
          line[pidx(pos)] = make_float2(v.x * t, v.y * t);
        }
        __builtin_amdgcn_wave_barrier();
        line_fft<MAXN>(line, a, lane, +1, tw);
      } else {
        line_fft<MAXN>(line, a, lane, a.sign, tw);
      }
    }
  }
  lds_barrier();

  // ---- store (with crop for adjoint passes) ----
  const bool crop = a.crop_out != 0;
  const int totout = lb * a.nout;
  if (ST == 2) {
    // every thread runs every round (the shuffles need the whole wave); a
    // group of lb adjacent lanes holds the lb coils of one output pixel
    const int64_t g = q0 >> lbl;
    float2* outp = a.out + (f * nib + g) * a.Ntot + (int64_t)a.N0 * rest;
    for (int b0 = 0; b0 < totout; b0 += kFftThreads * kUnroll) {
      float2 v[kUnroll];
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        const int idx = b0 + u * kFftThreads + tid;
        const int q = idx & (lb - 1), p = idx >> lbl;
        v[u] = make_float2(0.f, 0.f);
        if (idx < totout && q < nq) {
          int pos = p - a.nout / 2;
          if (pos < 0) pos += n;
          const float2 xv = lds[q * a.ldn + pidx(pos)];
          const float2 sv = a.S[f * a.Sfs + (p + (int64_t)a.N0 * rest) * a.C + q0 + q];
          v[u] = make_float2(sv.x * xv.x + sv.y * xv.y, sv.x * xv.y - sv.y * xv.x);
        }
      }
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        for (int m = lb >> 1; m > 0; m >>= 1) {
          v[u].x += __shfl_xor(v[u].x, m);
          v[u].y += __shfl_xor(v[u].y, m);
        }
        const int idx = b0 + u * kFftThreads + tid;
        if (idx < totout && (idx & (lb - 1)) == 0) outp[idx >> lbl] = v[u];
      }
    }
    return;
  }
  for (int b0 = tid; b0 < totout; b0 += kFftThreads * kUnroll) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int idx = b0 + u * kFftThreads;
      if (idx < totout) {
        const int q = idx & (lb - 1), p = idx >> lbl;
        if (q < nq) {
          int pos = p;
          if (crop) {
            pos = p - a.nout / 2;
            if (pos < 0) pos += n;
          }
          float2 v = lds[q * a.ldn + pidx(pos)];
          if (ST == 1) {
            const float dv = deapod(a, p, rest);
            v = make_float2(v.x * dv, v.y * dv);
          }
          a.out[(o * a.nout + p) * a.inner + q0 + q] = v;
        }
      }
    }
  }
}

size_t fft_lds_bytes(const FftArgs& a) { return sizeof(float2) * ((size_t)a.lb * a.ldn + a.n); }

void launch_fft(const FftArgs& a, int ld_mode, int st_mode, hipStream_t s) {
  if (a.use_reg) {
    launch_fft_reg(a, ld_mode, st_mode, s);
    return;
  }
  const int64_t nib = (a.inner + a.lb - 1) / a.lb;
  const dim3 grid((unsigned)(nib * a.outers));
  const size_t shm = fft_lds_bytes(a);
  const int conv = a.That != nullptr;
#define CGS_FFT_CASE(L, S_, CV)                                                          \
  if (ld_mode == L && st_mode == S_ && conv == CV) {                                     \
    if (a.n <= 512)                                                                      \
      CGS_LAUNCH((k_fft<L, S_, 512, CV>), grid, dim3(kFftThreads), shm, s, a);           \
    else                                                                                 \
      CGS_LAUNCH((k_fft<L, S_, 4096, CV>), grid, dim3(kFftThreads), shm, s, a);          \
    return;                                                                              \
  }
  CGS_FFT_CASE(0, 0, 0)
  CGS_FFT_CASE(1, 0, 0)
  CGS_FFT_CASE(2, 0, 0)
  CGS_FFT_CASE(3, 0, 0)
  CGS_FFT_CASE(4, 0, 0)
  CGS_FFT_CASE(0, 1, 0)
  CGS_FFT_CASE(0, 0, 1)
  CGS_FFT_CASE(0, 2, 0)
#undef CGS_FFT_CASE
}

}  // namespace cgs
