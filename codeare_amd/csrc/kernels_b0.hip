// kernels_b0.hip -- off-resonance (B0) as a third NFFT dimension, the
// inhomogeneity model of codeare's NFFT with "b0" and "timing"
// (NFFT.hpp:117-173, the mri_inh_3d plan of NFFTTraits.hpp:91-108,150-166,188-216):
//
//   f_j = sum_r x(r) exp(-2 pi i k_j.r) exp(-2 pi i b0(r) t_j)
//
// With ts the mid time, nu_r = N3 b0(r) / omega and tau_j = (t_j - ts) omega / N3
// (so that b0(r) (t_j - ts) = nu_r tau_j, |tau_j| <= 1 / (2 sigma3)), the
// Kaiser-Bessel pair phi(u) (support |u| < m) / phihat(tau) = I0(m sqrt(b3^2 -
// (2 pi tau)^2)), b3 = pi (2 - 1/sigma3), gives
//   sum_l phi(l - nu) exp(-2 pi i l tau) ~= phihat(tau) exp(-2 pi i nu tau),
// so f_j = NFFT3D[X](k_j, tau_j) / phihat(tau_j) with the 3D image
// X(r, l) = x(r) exp(-2 pi i b0(r) ts) phi(l - nu_r), l in [-N3/2, N3/2).
// spread builds X, gather applies the adjoint of that step, scale the
// 1/phihat(tau_j) of every sample.
#include <algorithm>

#include "cgs_internal.hpp"

namespace cgs {

// Kaiser-Bessel window of the third axis, truncated to |u| < m
__device__ __forceinline__ float b0_phi(float u, int m, float b) {
  const float s2 = (float)(m * m) - u * u;
  if (!(s2 > 0.f)) return 0.f;
  const float s = sqrtf(s2);
  return sinhf(b * s) / (3.14159265358979f * s);
}

// X3[((l * Ntot) + i) * C + c] = (mask) x[i * C + c] * phase[i] * phi(l - N3/2 - nu[i]);
// every element of X3 is written (zero outside the window)
__global__ void k_b0_spread(const float2* __restrict__ x, const float2* __restrict__ phase,
                            const float* __restrict__ nu, float2* __restrict__ X3, int64_t Ntot, int C, int N3,
                            int m, float b3, int mask_nh, int N0, int N1) {
  const int64_t tot = (int64_t)N3 * Ntot * C;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < tot; idx += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(idx % C);
    const int64_t li = idx / C;
    const int64_t i = li % Ntot;
    const int l = (int)(li / Ntot) - N3 / 2;
    const float u = (float)l - nu[i];
    float2 v = make_float2(0.f, 0.f);
    if (fabsf(u) < (float)m) {
      float wgt = b0_phi(u, m, b3);
      if (mask_nh > 0) {   // the CGNR's damping mask w_hat (disk of radius N0/2, NFFTTraits.hpp:229-256)
        const int64_t i0 = i % N0, i1 = i / N0;
        if ((i0 - mask_nh) * (i0 - mask_nh) + (i1 - mask_nh) * (i1 - mask_nh) > (int64_t)mask_nh * mask_nh)
          wgt = 0.f;
      }
      const float2 a = x[i * C + c], p = phase[i];
      v = make_float2(wgt * (a.x * p.x - a.y * p.y), wgt * (a.x * p.y + a.y * p.x));
    }
    X3[idx] = v;
  }
}

// x[i * C + c] = conj(phase[i]) * sum_l phi(l - N3/2 - nu[i]) X3[((l * Ntot) + i) * C + c]
__global__ void k_b0_gather(const float2* __restrict__ X3, const float2* __restrict__ phase,
                            const float* __restrict__ nu, float2* __restrict__ x, int64_t Ntot, int C, int N3, int m,
                            float b3) {
  const int64_t tot = Ntot * C;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < tot; idx += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(idx % C);
    const int64_t i = idx / C;
    const float v = nu[i];
    const int lo = max(-N3 / 2, (int)floorf(v) - m), hi = min(N3 / 2 - 1, (int)floorf(v) + m + 1);
    float sx = 0.f, sy = 0.f;
    for (int l = lo; l <= hi; ++l) {
      const float u = (float)l - v;
      if (fabsf(u) < (float)m) {
        const float wgt = b0_phi(u, m, b3);
        const float2 a = X3[((int64_t)(l + N3 / 2) * Ntot + i) * C + c];
        sx += wgt * a.x;
        sy += wgt * a.y;
      }
    }
    const float2 p = phase[i];
    x[idx] = make_float2(sx * p.x + sy * p.y, sy * p.x - sx * p.y);
  }
}

// y[s * C + c] *= scale[s] (in place, or into out)
__global__ void k_b0_scale(const float2* __restrict__ y, const float* __restrict__ scale, float2* __restrict__ out,
                           int64_t M, int C) {
  const int64_t tot = M * C;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < tot; idx += (int64_t)gridDim.x * blockDim.x) {
    const float s = scale[idx / C];
    const float2 v = y[idx];
    out[idx] = make_float2(v.x * s, v.y * s);
  }
}

static unsigned blocks_for(int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192));
}

void launch_b0_spread(const float2* x, const float2* phase, const float* nu, float2* X3, int64_t Ntot, int C, int N3,
                      int m, float b3, int mask_nh, int N0, int N1, hipStream_t s) {
  hipLaunchKernelGGL(k_b0_spread, dim3(blocks_for((int64_t)N3 * Ntot * C)), dim3(256), 0, s, x, phase, nu, X3, Ntot,
                     C, N3, m, b3, mask_nh, N0, N1);
}

void launch_b0_gather(const float2* X3, const float2* phase, const float* nu, float2* x, int64_t Ntot, int C, int N3,
                      int m, float b3, hipStream_t s) {
  hipLaunchKernelGGL(k_b0_gather, dim3(blocks_for(Ntot * C)), dim3(256), 0, s, X3, phase, nu, x, Ntot, C, N3, m, b3);
}

void launch_b0_scale(const float2* y, const float* scale, float2* out, int64_t M, int C, hipStream_t s) {
  hipLaunchKernelGGL(k_b0_scale, dim3(blocks_for(M * C)), dim3(256), 0, s, y, scale, out, M, C);
}

}  // namespace cgs
