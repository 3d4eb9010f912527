// kernels_vec.hip -- coil combine, CG / CGNR vector and scalar kernels, layout
// conversions.
//
// Reductions are deterministic: every block writes an FP64 partial to a fixed
// slot and a one-block kernel per frame sums the partials in a fixed tree order.
// The CG scalars (rn, <p,q>, alpha, beta, residual) live on the device in FP64
// (CgState); vector arithmetic stays FP32.  Every kernel skips frames whose
// done flag is set, so finished frames cost only a flag read.
#include <algorithm>

#include "cgs_internal.hpp"

namespace cgs {


__device__ __forceinline__ double block_sum(double v, double* sh) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  v = wave_sum_d(v);
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int k = 0; k < nw; ++k) t += sh[k];
  }
  __syncthreads();
  return t;   // valid in thread 0
}

// part[t], part[t + blockDim], ... added in that order by thread t, eight loads
// issued before they are added (the order, and so the sum, is that of a plain loop)
__device__ __forceinline__ double strided_sum(const double* part, int nb) {
  double acc = 0.0;
  const int st = (int)blockDim.x;
  int b = threadIdx.x;
  for (; b + 7 * st < nb; b += 8 * st) {
    double t[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) t[u] = part[b + u * st];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc += t[u];
  }
  for (; b < nb; b += st) acc += part[b];
  return acc;
}


// ---------------------------------------------------------------------------
// x[f][i] = (sum_c conj(S[i][c]) X[f][i][c]) * I[i] + lambda p[f][i]; partial <p, x>
// CP lanes per pixel (power of two >= C, at most 64).
template <int CP>
__global__ void __launch_bounds__(256) k_coil_combine(const float2* X, const float2* S,
                                                      const float* Imap, float2* x,
                                                      const float2* pv, float lambda,
                                                      double* partials, int C, int64_t Ntot,
                                                      const int* fdone, int64_t Sfs, int64_t Ifs) {
  __shared__ double sh[4];
  const int f = blockIdx.y;
  if (fdone && fdone[f]) return;
  S += f * Sfs;
  if (Imap) Imap += f * Ifs;
  constexpr int PPB = 256 / CP;
  const int g = threadIdx.x / CP, cl = threadIdx.x % CP;
  double dot = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * PPB + g; i < Ntot; i += (int64_t)gridDim.x * PPB) {
    float2 acc = make_float2(0.f, 0.f);
    for (int c = cl; c < C; c += CP) {
      const float2 s = S[i * C + c];
      const float2 v = X[((int64_t)f * Ntot + i) * C + c];
      acc.x += s.x * v.x + s.y * v.y;
      acc.y += s.x * v.y - s.y * v.x;
    }
    acc = make_float2(group_sum<CP>(acc.x), group_sum<CP>(acc.y));   // over the pixel's CP lanes
    if (cl == 0) {
      if (Imap) { const float iv = Imap[i]; acc.x *= iv; acc.y *= iv; }
      const int64_t o = (int64_t)f * Ntot + i;
      if (pv) {
        const float2 p = pv[o];
        acc.x += lambda * p.x;
        acc.y += lambda * p.y;
        dot += (double)p.x * acc.x + (double)p.y * acc.y;
      }
      x[o] = acc;
    }
  }
  if (partials) {
    const double t = block_sum(dot, sh);
    if (threadIdx.x == 0) partials[(int64_t)f * gridDim.x + blockIdx.x] = t;
  }
}

void launch_coil_combine(const Plan& p, const float2* Xc, float2* x, const float2* pvec,
                         float lambda, bool use_imap, double* partials, const int* fl, hipStream_t s) {
  // kCoilBlocks partials per frame (read by k_cg_alpha and k_cg_xr_alpha)
  const dim3 g(kCoilBlocks, p.F);
  const float* I = use_imap ? p.Imap : nullptr;
  if (p.C <= 8)
    hipLaunchKernelGGL(k_coil_combine<8>, g, dim3(256), 0, s, Xc, p.S, I, x, pvec, lambda, partials, p.C, p.Ntot, fl,
                       p.Sfs, p.Ifs);
  else if (p.C <= 16)
    hipLaunchKernelGGL(k_coil_combine<16>, g, dim3(256), 0, s, Xc, p.S, I, x, pvec, lambda, partials, p.C, p.Ntot, fl,
                       p.Sfs, p.Ifs);
  else if (p.C <= 32)
    hipLaunchKernelGGL(k_coil_combine<32>, g, dim3(256), 0, s, Xc, p.S, I, x, pvec, lambda, partials, p.C, p.Ntot, fl,
                       p.Sfs, p.Ifs);
  else
    hipLaunchKernelGGL(k_coil_combine<64>, g, dim3(256), 0, s, Xc, p.S, I, x, pvec, lambda, partials, p.C, p.Ntot, fl,
                       p.Sfs, p.Ifs);
}

// q = sum_g part_g + lambda p, with <p, q> partials (kCoilBlocks per frame,
// the layout k_coil_combine writes); part[(f*G + g)*Ntot + i] are the partial
// coil sums of the fused crop pass
__global__ void __launch_bounds__(256) k_partial_combine(const float2* __restrict__ part, int G, int64_t Ntot,
                                                         float2* __restrict__ q, const float2* __restrict__ pv,
                                                         float lambda, double* __restrict__ partials,
                                                         const int* __restrict__ fdone) {
  __shared__ double sh[4];
  const int f = blockIdx.y;
  if (fdone && fdone[f]) return;
  double dot = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < Ntot; i += (int64_t)gridDim.x * 256) {
    float2 acc = make_float2(0.f, 0.f);
    for (int g = 0; g < G; ++g) {
      const float2 v = part[((int64_t)f * G + g) * Ntot + i];
      acc.x += v.x;
      acc.y += v.y;
    }
    const int64_t o = (int64_t)f * Ntot + i;
    if (pv) {
      const float2 p = pv[o];
      acc.x += lambda * p.x;
      acc.y += lambda * p.y;
      dot += (double)p.x * acc.x + (double)p.y * acc.y;
    }
    q[o] = acc;
  }
  if (partials) {
    const double t = block_sum(dot, sh);
    if (threadIdx.x == 0) partials[(int64_t)f * gridDim.x + blockIdx.x] = t;
  }
}

// Fused CG tail of the Toeplitz path: alpha = rn / (sum pq_part + lambda sum
// pp_part) (every block sums the same partials in the same order), then
// q = sum_g part_g + lambda p computed in place, x += alpha p, r -= alpha q and
// ||r||^2 partials.  The first pixel of every thread is loaded before the
// partial sums so that both latencies overlap.  snap: record rn and it for the
// pad pass that forms beta (CgState::bpend)
__global__ void __launch_bounds__(256) k_cg_tail(float2* x, float2* r, const float2* pv, const float2* part, int G,
                                                 float lambda, CgState* st, const double* pq_part, int npq,
                                                 const double* pp_part, int npp, double* partials, int64_t Ntot,
                                                 const int* fdone, int snap) {
  __shared__ double sh[4], sh2[4];
  __shared__ float alpha_s;
  const int f = blockIdx.y;
  if (fdone[f]) return;
  const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool has = i0 < Ntot;
  float2 p0 = make_float2(0.f, 0.f), q0 = p0, x0 = p0, r0 = p0;
  if (has) {
    const int64_t o = (int64_t)f * Ntot + i0;
    p0 = pv[o];
    q0 = make_float2(lambda * p0.x, lambda * p0.y);
    for (int g = 0; g < G; ++g) {
      const float2 v = part[((int64_t)f * G + g) * Ntot + i0];
      q0.x += v.x;
      q0.y += v.y;
    }
    x0 = x[o];
    r0 = r[o];
  }
  {
    // every block sums the partials in the same order; four loads in flight
    // per thread, then one LDS round (blockDim = 256)
    const double* pqf = pq_part + (int64_t)f * npq;
    const double* ppf = pp_part + (int64_t)f * npp;
    auto sum4 = [](const double* v, int n) {   // v[t], v[t + 256], ... with 4 loads issued together
      double acc = 0.0;
      for (int i = threadIdx.x; i < n; i += 1024) {
        const double t0 = v[i];
        const double t1 = i + 256 < n ? v[i + 256] : 0.0;
        const double t2 = i + 512 < n ? v[i + 512] : 0.0;
        const double t3 = i + 768 < n ? v[i + 768] : 0.0;
        acc += (t0 + t1) + (t2 + t3);
      }
      return acc;
    };
    double a = sum4(pqf, npq), b = sum4(ppf, npp);
    a = wave_sum_d(a);
    b = wave_sum_d(b);
    if ((threadIdx.x & 63) == 0) {
      sh[threadIdx.x >> 6] = a;
      sh2[threadIdx.x >> 6] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      const double pq = ((sh[0] + sh[1]) + (sh[2] + sh[3])) + (double)lambda * ((sh2[0] + sh2[1]) + (sh2[2] + sh2[3]));
      const double al = st[f].rn / pq;   // CGLS.hpp:50
      alpha_s = (float)al;
      if (blockIdx.x == 0) {
        st[f].pq = pq;
        st[f].alpha = al;
        if (snap) {
          st[f].rn_snap = st[f].rn;
          st[f].it_snap = st[f].it;
          st[f].bpend = 1;
        }
      }
    }
  }
  __syncthreads();
  const float al = alpha_s;
  double acc = 0.0;
  if (has) {
    const int64_t o = (int64_t)f * Ntot + i0;
    x0.x += al * p0.x; x0.y += al * p0.y;
    r0.x -= al * q0.x; r0.y -= al * q0.y;
    x[o] = x0;
    r[o] = r0;
    acc += (double)r0.x * r0.x + (double)r0.y * r0.y;
  }
  for (int64_t i = i0 + (int64_t)gridDim.x * blockDim.x; i < Ntot; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t o = (int64_t)f * Ntot + i;
    const float2 pp = pv[o];
    float2 qq = make_float2(lambda * pp.x, lambda * pp.y);
    for (int g = 0; g < G; ++g) {
      const float2 v = part[((int64_t)f * G + g) * Ntot + i];
      qq.x += v.x;
      qq.y += v.y;
    }
    float2 xx = x[o], rr = r[o];
    xx.x += al * pp.x; xx.y += al * pp.y;
    rr.x -= al * qq.x; rr.y -= al * qq.y;
    x[o] = xx;
    r[o] = rr;
    acc += (double)rr.x * rr.x + (double)rr.y * rr.y;
  }
  const double t = block_sum(acc, sh);
  if (threadIdx.x == 0) partials[(int64_t)f * gridDim.x + blockIdx.x] = t;
}

// ||r||^2 partials of the CG tail: cg_tail_blocks(p) per frame (one pixel
// per thread up to kRedBlocks blocks)
const double* cg_tail_rpartials(const Plan& p) { return p.partials + (size_t)kCoilBlocks * p.F; }
int cg_tail_blocks(const Plan& p) {
  return (int)std::max<int64_t>(1, std::min<int64_t>(kRedBlocks, (p.Ntot + 255) / 256));
}

void launch_cg_tail(const Plan& p, float2* x, float2* r, const float2* pv, const float2* part, int G,
                    float lambda, const double* pq_part, int npq, const double* pp_part, int npp, bool snap,
                    hipStream_t s) {
  hipLaunchKernelGGL(k_cg_tail, dim3(cg_tail_blocks(p), p.F), dim3(256), 0, s, x, r, pv, part, G, lambda, p.cgst,
                     pq_part, npq, pp_part, npp, p.partials + (size_t)kCoilBlocks * p.F, p.Ntot, p.fdone,
                     snap ? 1 : 0);
}

void launch_partial_combine(const Plan& p, const float2* part, int G, float2* q, const float2* pvec, float lambda,
                            double* partials, const int* fl, hipStream_t s) {
  const dim3 g(kCoilBlocks, p.F);
  hipLaunchKernelGGL(k_partial_combine, g, dim3(256), 0, s, part, G, p.Ntot, q, pvec, lambda, partials, fl);
}

// ---------------------------------------------------------------------------
// Fused CG iteration tail.  Blocks of a frame write FP64 partials; the last
// block to arrive (global ticket per frame, reset by that block) sums them in
// fixed order and updates the frame's scalars, so no separate scalar kernel is
// launched and the result does not depend on block timing.
__device__ __forceinline__ bool last_block(int* ticket, int nblocks) {
  __shared__ int is_last;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    const int t = atomicAdd(ticket, 1);
    is_last = (t == nblocks - 1);
    if (is_last) atomicExch(ticket, 0);
  }
  __syncthreads();
  if (is_last) __threadfence();
  return is_last != 0;
}

// q = sum_c conj(S_c) X_c + lambda p, partial <p, q>; last block: alpha = rn / <p,q>
template <int CP>
__global__ void __launch_bounds__(256) k_coil_alpha(const float2* X, const float2* S, float2* q,
                                                    const float2* pv, float lambda, double* partials,
                                                    int* tickets, CgState* st, int C, int64_t Ntot,
                                                    const int* fdone, int64_t Sfs) {
  __shared__ double sh[4];
  const int f = blockIdx.y;
  if (fdone[f]) return;
  S += f * Sfs;
  constexpr int PPB = 256 / CP;
  const int g = threadIdx.x / CP, cl = threadIdx.x % CP;
  double dot = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * PPB + g; i < Ntot; i += (int64_t)gridDim.x * PPB) {
    float2 acc = make_float2(0.f, 0.f);
    for (int c = cl; c < C; c += CP) {
      const float2 s = S[i * C + c];
      const float2 v = X[((int64_t)f * Ntot + i) * C + c];
      acc.x += s.x * v.x + s.y * v.y;
      acc.y += s.x * v.y - s.y * v.x;
    }
    acc = make_float2(group_sum<CP>(acc.x), group_sum<CP>(acc.y));   // over the pixel's CP lanes
    if (cl == 0) {
      const int64_t o = (int64_t)f * Ntot + i;
      const float2 p = pv[o];
      acc.x += lambda * p.x;
      acc.y += lambda * p.y;
      dot += (double)p.x * acc.x + (double)p.y * acc.y;
      q[o] = acc;
    }
  }
  const double t = block_sum(dot, sh);
  if (threadIdx.x == 0) partials[(int64_t)f * gridDim.x + blockIdx.x] = t;
  if (last_block(tickets + f, gridDim.x)) {
    double acc = 0.0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x) acc += partials[(int64_t)f * gridDim.x + b];
    const double pq = block_sum(acc, sh);
    if (threadIdx.x == 0) {
      st[f].pq = pq;
      st[f].alpha = st[f].rn / pq;   // CGLS.hpp:50
    }
  }
}

void launch_coil_alpha(const Plan& p, const float2* Xc, float2* q, const float2* pv, float lambda,
                       hipStream_t s) {
  const dim3 g(kRedBlocks, p.F);
  if (p.C <= 8)
    hipLaunchKernelGGL(k_coil_alpha<8>, g, dim3(256), 0, s, Xc, p.S, q, pv, lambda, p.partials, p.tickets, p.cgst, p.C, p.Ntot, p.fdone,
                       p.Sfs);
  else if (p.C <= 16)
    hipLaunchKernelGGL(k_coil_alpha<16>, g, dim3(256), 0, s, Xc, p.S, q, pv, lambda, p.partials, p.tickets, p.cgst, p.C, p.Ntot, p.fdone,
                       p.Sfs);
  else if (p.C <= 32)
    hipLaunchKernelGGL(k_coil_alpha<32>, g, dim3(256), 0, s, Xc, p.S, q, pv, lambda, p.partials, p.tickets, p.cgst, p.C, p.Ntot, p.fdone,
                       p.Sfs);
  else
    hipLaunchKernelGGL(k_coil_alpha<64>, g, dim3(256), 0, s, Xc, p.S, q, pv, lambda, p.partials, p.tickets, p.cgst, p.C, p.Ntot, p.fdone,
                       p.Sfs);
}

// x += alpha p ; r -= alpha q ; partial ||r||^2, where every block forms
// alpha = rn / <p,q> itself from the coil combine's partials (same fixed-order
// sum in every block, so all blocks use the identical alpha)
__global__ void __launch_bounds__(256) k_cg_xr_alpha(float2* x, float2* r, const float2* pv, const float2* q,
                                                     CgState* st, const double* pq_part, int npq, double* partials,
                                                     int64_t Ntot, const int* fdone) {
  __shared__ double sh[4];
  __shared__ float alpha_s;
  const int f = blockIdx.y;
  if (fdone[f]) return;
  {
    const double pq = block_sum(strided_sum(pq_part + (int64_t)f * npq, npq), sh);
    if (threadIdx.x == 0) {
      const double al = st[f].rn / pq;   // CGLS.hpp:50
      alpha_s = (float)al;
      if (blockIdx.x == 0) { st[f].pq = pq; st[f].alpha = al; }
    }
    __syncthreads();
  }
  const float al = alpha_s;
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < Ntot; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t o = (int64_t)f * Ntot + i;
    const float2 pp = pv[o], qq = q[o];
    float2 xx = x[o], rr = r[o];
    xx.x += al * pp.x; xx.y += al * pp.y;
    rr.x -= al * qq.x; rr.y -= al * qq.y;
    x[o] = xx;
    r[o] = rr;
    acc += (double)rr.x * rr.x + (double)rr.y * rr.y;
  }
  const double t = block_sum(acc, sh);
  if (threadIdx.x == 0) partials[(int64_t)f * gridDim.x + blockIdx.x] = t;
}

void launch_cg_xr_alpha(const Plan& p, float2* x, float2* r, const float2* pv, const float2* q, hipStream_t s) {
  hipLaunchKernelGGL(k_cg_xr_alpha, dim3(kRedBlocks, p.F), dim3(256), 0, s, x, r, pv, q, p.cgst, p.partials,
                     kCoilBlocks, p.partials + (size_t)kCoilBlocks * p.F, p.Ntot, p.fdone);
}

__global__ void k_cg_beta(CgState* st, const double* part, int nb, int tolkind, double tol, int maxit,
                          double* hist, int hcap, int* fdone, int nf);

void launch_cg_beta_after_xr(const Plan& p, int nparts, int tolkind, double tol, int maxit, hipStream_t s) {
  hipLaunchKernelGGL(k_cg_beta, dim3(p.F), dim3(256), 0, s, p.cgst, p.partials + (size_t)kCoilBlocks * p.F, nparts,
                     tolkind, tol, maxit, p.hist, p.hist_cap, p.fdone, 1);
}

// x += alpha p ; r -= alpha q ; partial ||r||^2 ; last block: beta, residual, done
__global__ void k_cg_xr_beta(float2* x, float2* r, const float2* pv, const float2* q, CgState* st,
                             double* partials, int* tickets, int64_t Ntot, int tolkind, double tol, int maxit,
                             double* hist, int hcap, int* fdone) {
  __shared__ double sh[4];
  const int f = blockIdx.y;
  if (fdone[f]) return;
  const float al = (float)st[f].alpha;
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < Ntot; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t o = (int64_t)f * Ntot + i;
    const float2 pp = pv[o], qq = q[o];
    float2 xx = x[o], rr = r[o];
    xx.x += al * pp.x; xx.y += al * pp.y;
    rr.x -= al * qq.x; rr.y -= al * qq.y;
    x[o] = xx;
    r[o] = rr;
    acc += (double)rr.x * rr.x + (double)rr.y * rr.y;
  }
  const double t = block_sum(acc, sh);
  if (threadIdx.x == 0) partials[(int64_t)f * gridDim.x + blockIdx.x] = t;
  if (last_block(tickets + f, gridDim.x)) {
    double a2 = 0.0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x) a2 += partials[(int64_t)f * gridDim.x + b];
    const double rn_new = block_sum(a2, sh);
    if (threadIdx.x == 0) {
      CgState c = st[f];
      c.beta = rn_new / c.rn;   // CGLS.hpp:53-55
      c.rn = rn_new;
      c.it += 1;
      const double rel = (tolkind == CGS_TOL_SQUARED_RATIO) ? rn_new / c.xn : sqrt(rn_new / c.bn);
      c.rel = rel;
      if (isnan(rel) || isnan(c.alpha)) c.nan = 1;
      c.done = (c.nan || rel <= tol || c.it >= maxit) ? 1 : 0;
      st[f] = c;
      if (hist && c.it < hcap) hist[(int64_t)f * hcap + c.it] = rel;
      fdone[f] = c.done;
    }
  }
}

void launch_cg_xr_beta(const Plan& p, float2* x, float2* r, const float2* pv, const float2* q, int tolkind,
                       double tol, int maxit, hipStream_t s) {
  hipLaunchKernelGGL(k_cg_xr_beta, dim3(kRedBlocks, p.F), dim3(256), 0, s, x, r, pv, q, p.cgst,
                     p.partials + (size_t)kRedBlocks * p.F, p.tickets + p.F, p.Ntot, tolkind, tol, maxit, p.hist,
                     p.hist_cap, p.fdone);
}

// ---------------------------------------------------------------------------
__global__ void k_sqnorm_partials(const float2* v, int64_t Ntot, double* partials, const int* fdone) {
  __shared__ double sh[4];
  const int f = blockIdx.y;
  if (fdone && fdone[f]) return;
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < Ntot; i += (int64_t)gridDim.x * blockDim.x) {
    const float2 a = v[(int64_t)f * Ntot + i];
    acc += (double)a.x * a.x + (double)a.y * a.y;
  }
  const double t = block_sum(acc, sh);
  if (threadIdx.x == 0) partials[(int64_t)f * gridDim.x + blockIdx.x] = t;
}

void launch_sqnorm_partials(const Plan& p, const float2* v, double* partials, hipStream_t s) {
  hipLaunchKernelGGL(k_sqnorm_partials, dim3(kRedBlocks, p.F), dim3(256), 0, s, v, p.Ntot, partials, p.fdone);
}

__device__ double sum_partials(const double* part, int nb, double* sh) {
  return block_sum(strided_sum(part, nb), sh);
}

// The scalar kernels below run one block per frame, or, for a joint CG over
// all frames (nf = F, one block), sum the partials of every frame (frame-major,
// a fixed order) and give every frame the same scalars and stop flag.

// CG start: bn = ||E^H y||^2, rn = ||r0||^2, residual measure, done test.
__global__ void k_cg_init(CgState* st, const double* pbn, const double* prn, int nb, int tolkind,
                          double tol, int maxit, double* hist, int hcap, int* fdone, int nf) {
  __shared__ double sh[4];
  const int f = blockIdx.x;
  const double bn = sum_partials(pbn + (int64_t)f * nb, nb * nf, sh);
  const double rn = prn ? sum_partials(prn + (int64_t)f * nb, nb * nf, sh) : bn;
  if (threadIdx.x == 0) {
    CgState c;
    c.bn = bn;
    c.rn = rn;
    c.xn = rn;   // codeare: rn = xn = Re<r, p> (CGLS.hpp:35)
    c.pq = 0; c.alpha = 0; c.beta = 0;
    c.it = 0; c.nan = 0; c.pad = 0;
    c.rn_snap = 0; c.it_snap = 0; c.bpend = 0;
    double rel;
    if (tolkind == CGS_TOL_SQUARED_RATIO) { c.xn = bn; rel = bn > 0 ? rn / bn : 0.0; }
    else rel = bn > 0 ? sqrt(rn / bn) : 0.0;
    // a NaN in E^H y (bn) or in r0 stops the frame with the NaN flag, as
    // codeare's isnan(rn/xn) test does (CGLS.hpp:41); bn == 0 is a zero
    // right-hand side, solved by x = x0
    if (isnan(bn) || isnan(rn)) rel = bn + rn;
    c.rel = rel;
    c.nan = isnan(rel) ? 1 : 0;
    c.done = (c.nan || rel <= tol || maxit <= 0) ? 1 : 0;
    for (int g = f; g < f + nf; ++g) {
      st[g] = c;
      fdone[g] = c.done;
      if (hist) hist[(int64_t)g * hcap] = rel;
    }
  }
}

void launch_cg_init(const Plan& p, const double* part_bn, const double* part_rn, int tolkind,
                    double tol, int maxit, bool warm, bool joint, hipStream_t s) {
  hipLaunchKernelGGL(k_cg_init, dim3(joint ? 1 : p.F), dim3(256), 0, s, p.cgst, part_bn, warm ? part_rn : nullptr,
                     kRedBlocks, tolkind, tol, maxit, p.hist, p.hist_cap, p.fdone, joint ? p.F : 1);
}

__global__ void k_cg_alpha(CgState* st, const double* part, int nb, const int* fdone, int nf) {
  __shared__ double sh[4];
  const int f = blockIdx.x;
  if (fdone[f]) return;
  const double pq = sum_partials(part + (int64_t)f * nb, nb * nf, sh);
  if (threadIdx.x == 0) {
    const double al = st[f].rn / pq;   // CGLS.hpp:50
    for (int g = f; g < f + nf; ++g) {
      st[g].pq = pq;
      st[g].alpha = al;
    }
  }
}

void launch_cg_alpha(const Plan& p, const double* partials, bool joint, hipStream_t s) {
  hipLaunchKernelGGL(k_cg_alpha, dim3(joint ? 1 : p.F), dim3(256), 0, s, p.cgst, partials, kCoilBlocks, p.fdone,
                     joint ? p.F : 1);
}

// x += alpha p ; r -= alpha q ; partial ||r||^2   (CGLS.hpp:51-54)
__global__ void k_cg_xr(float2* x, float2* r, const float2* pv, const float2* q, const CgState* st,
                        double* partials, int64_t Ntot, const int* fdone) {
  __shared__ double sh[4];
  const int f = blockIdx.y;
  if (fdone[f]) return;
  const float al = (float)st[f].alpha;
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < Ntot; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t o = (int64_t)f * Ntot + i;
    const float2 pp = pv[o], qq = q[o];
    float2 xx = x[o], rr = r[o];
    xx.x += al * pp.x; xx.y += al * pp.y;
    rr.x -= al * qq.x; rr.y -= al * qq.y;
    x[o] = xx;
    r[o] = rr;
    acc += (double)rr.x * rr.x + (double)rr.y * rr.y;
  }
  const double t = block_sum(acc, sh);
  if (threadIdx.x == 0) partials[(int64_t)f * gridDim.x + blockIdx.x] = t;
}

void launch_cg_xr(const Plan& p, float2* x, float2* r, const float2* pv, const float2* q,
                  double* partials, hipStream_t s) {
  hipLaunchKernelGGL(k_cg_xr, dim3(kRedBlocks, p.F), dim3(256), 0, s, x, r, pv, q, p.cgst, partials, p.Ntot, p.fdone);
}

__global__ void k_cg_beta(CgState* st, const double* part, int nb, int tolkind, double tol, int maxit,
                          double* hist, int hcap, int* fdone, int nf) {
  __shared__ double sh[4];
  const int f = blockIdx.x;
  if (fdone[f]) return;
  const double rn_new = sum_partials(part + (int64_t)f * nb, nb * nf, sh);
  if (threadIdx.x == 0) {
    CgState c = st[f];
    c.beta = rn_new / c.rn;   // CGLS.hpp:53-55
    c.rn = rn_new;
    c.it += 1;
    const double rel = (tolkind == CGS_TOL_SQUARED_RATIO) ? rn_new / c.xn : sqrt(rn_new / c.bn);
    c.rel = rel;
    if (isnan(rel) || isnan(c.alpha)) c.nan = 1;
    c.done = (c.nan || rel <= tol || c.it >= maxit) ? 1 : 0;
    for (int g = f; g < f + nf; ++g) {
      st[g] = c;
      if (hist && c.it < hcap) hist[(int64_t)g * hcap + c.it] = rel;
      fdone[g] = c.done;
    }
  }
}

void launch_cg_beta(const Plan& p, const double* partials, int tolkind, double tol, int maxit, bool joint,
                    hipStream_t s) {
  hipLaunchKernelGGL(k_cg_beta, dim3(joint ? 1 : p.F), dim3(256), 0, s, p.cgst, partials, kRedBlocks, tolkind, tol,
                     maxit, p.hist, p.hist_cap, p.fdone, joint ? p.F : 1);
}

// p = r + beta p  (CGLS.hpp:55-56)
__global__ void k_cg_p(float2* pv, const float2* r, const CgState* st, int64_t Ntot, const int* fdone) {
  const int f = blockIdx.y;
  if (fdone[f]) return;
  const float be = (float)st[f].beta;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < Ntot; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t o = (int64_t)f * Ntot + i;
    const float2 rr = r[o];
    float2 pp = pv[o];
    pp.x = rr.x + be * pp.x;
    pp.y = rr.y + be * pp.y;
    pv[o] = pp;
  }
}

void launch_cg_p(const Plan& p, float2* pv, const float2* r, hipStream_t s) {
  hipLaunchKernelGGL(k_cg_p, dim3(kRedBlocks, p.F), dim3(256), 0, s, pv, r, p.cgst, p.Ntot, p.fdone);
}

// out = alpha a + beta b over all frames (no done test)
__global__ void k_axpby(float2* out, const float2* a, float al, const float2* b, float be, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    float2 u = a ? a[i] : make_float2(0.f, 0.f);
    float2 v = b ? b[i] : make_float2(0.f, 0.f);
    out[i] = make_float2(al * u.x + be * v.x, al * u.y + be * v.y);
  }
}

void launch_axpby_img(const Plan& p, float2* out, const float2* a, float alpha, const float2* b,
                      float beta, hipStream_t s) {
  hipLaunchKernelGGL(k_axpby, dim3(1024), dim3(256), 0, s, out, a, alpha, b, beta, p.Ntot * p.F);
}

// ---------------------------------------------------------------------------
// Layout conversions between codeare's (M_f, C) frame blocks and the sorted,
// coil-interleaved internal sample layout.
__global__ void k_y_ext2int(const float2* ext, float2* in, const int* perm, const int* sframe,
                            const int64_t* foff, int C, int batch, int only_frame, int64_t Mtot) {
  const int64_t tot = Mtot * C;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < tot; idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = idx / C;
    const int c = (int)(idx - s * C);
    const int f = sframe[s];
    if (only_frame >= 0 && f != only_frame) continue;
    float2 v = make_float2(0.f, 0.f);
    if (c < batch) {
      const int64_t j = perm[s], o = foff[f], Mf = foff[f + 1] - o;
      const int64_t e = (only_frame >= 0 ? 0 : (int64_t)batch * o) + (j - o) + Mf * c;
      v = ext[e];
    }
    in[idx] = v;
  }
}

__global__ void k_y_int2ext(const float2* in, float2* ext, const int* perm, const int* sframe,
                            const int64_t* foff, int C, int batch, int only_frame, int64_t Mtot) {
  const int64_t tot = Mtot * C;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < tot; idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = idx / C;
    const int c = (int)(idx - s * C);
    if (c >= batch) continue;
    const int f = sframe[s];
    if (only_frame >= 0 && f != only_frame) continue;
    const int64_t j = perm[s], o = foff[f], Mf = foff[f + 1] - o;
    const int64_t e = (only_frame >= 0 ? 0 : (int64_t)batch * o) + (j - o) + Mf * c;
    ext[e] = in[idx];
  }
}

// The same conversions over all frames as a transpose through LDS: a block
// takes 64 consecutive samples of codeare's order, reads or writes each one's
// C coils as one row of the internal layout (C x 8 contiguous bytes) and each
// coil's 64 samples as one run of codeare's column (512 contiguous bytes).
constexpr int kYT = 64;   // samples per block
// per sample of the block (codeare order j = j0 + jj, j < jend): its internal
// index, its column base in codeare's layout and its frame's sample count (the
// column stride).  all_frames: codeare's frame blocks one after the other
// (batch columns each); else the one frame's block alone.
__device__ __forceinline__ void y_tile_index(const int* iperm, const int* sframe, const int64_t* foff, int batch,
                                             int64_t j0, int64_t jend, bool all_frames, int* si, int64_t* eb,
                                             int64_t* mf) {
  const int jj = threadIdx.x;
  if (jj < kYT) {
    const int64_t j = j0 + jj;
    if (j < jend) {
      const int s = iperm[j];
      const int f = sframe[s];
      const int64_t o = foff[f];
      si[jj] = s;
      eb[jj] = (all_frames ? (int64_t)batch * o : 0) + (j - o);
      mf[jj] = foff[f + 1] - o;
    }
  }
  __syncthreads();
}

// blocks cover codeare's samples [jbeg, jend): all of them (only_frame < 0) or
// one frame's
__global__ void __launch_bounds__(256) k_y_ext2int_t(const float2* ext, float2* in, const int* iperm,
                                                     const int* sframe, const int64_t* foff, int C, int batch,
                                                     int64_t jbeg, int64_t jend, int only_frame) {
  extern __shared__ float2 yt[];                 // [kYT][C + 1]
  __shared__ int si[kYT];
  __shared__ int64_t eb[kYT], mf[kYT];
  const int64_t j0 = jbeg + (int64_t)blockIdx.x * kYT;
  const int nj = (int)min<int64_t>(kYT, jend - j0);
  const int ld = C + 1;
  y_tile_index(iperm, sframe, foff, batch, j0, jend, only_frame < 0, si, eb, mf);
  for (int e = threadIdx.x; e < kYT * batch; e += blockDim.x) {   // coil-major: runs of codeare's column
    const int c = e / kYT, jj = e - c * kYT;
    if (jj < nj) yt[jj * ld + c] = ext[eb[jj] + mf[jj] * c];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < kYT * C; e += blockDim.x) {       // sample-major: rows of the internal layout
    const int jj = e / C, c = e - jj * C;
    if (jj < nj) in[(int64_t)si[jj] * C + c] = c < batch ? yt[jj * ld + c] : make_float2(0.f, 0.f);
  }
}

__global__ void __launch_bounds__(256) k_y_int2ext_t(const float2* in, float2* ext, const int* iperm,
                                                     const int* sframe, const int64_t* foff, int C, int batch,
                                                     int64_t jbeg, int64_t jend, int only_frame) {
  extern __shared__ float2 yt[];                 // [kYT][C + 1]
  __shared__ int si[kYT];
  __shared__ int64_t eb[kYT], mf[kYT];
  const int64_t j0 = jbeg + (int64_t)blockIdx.x * kYT;
  const int nj = (int)min<int64_t>(kYT, jend - j0);
  const int ld = C + 1;
  y_tile_index(iperm, sframe, foff, batch, j0, jend, only_frame < 0, si, eb, mf);
  for (int e = threadIdx.x; e < kYT * batch; e += blockDim.x) {   // sample-major reads
    const int jj = e / batch, c = e - jj * batch;
    if (jj < nj) yt[jj * ld + c] = in[(int64_t)si[jj] * C + c];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < kYT * batch; e += blockDim.x) {   // coil-major writes
    const int c = e / kYT, jj = e - c * kYT;
    if (jj < nj) ext[eb[jj] + mf[jj] * c] = yt[jj * ld + c];
  }
}

// the samples of codeare's order a conversion covers
static void y_range(const Plan& p, int only_frame, int64_t* jbeg, int64_t* jend) {
  *jbeg = only_frame < 0 ? 0 : p.foff[only_frame];
  *jend = only_frame < 0 ? p.Mtot : p.foff[only_frame + 1];
}
void launch_y_ext2int(const Plan& p, const float2* ext, float2* in, int batch, int only_frame, hipStream_t s) {
  if (p.iperm && batch <= p.C && p.C < 128) {   // its LDS tile within 64 KB
    int64_t jb, je;
    y_range(p, only_frame, &jb, &je);
    if (je <= jb) return;
    const size_t shm = sizeof(float2) * kYT * (size_t)(p.C + 1);
    hipLaunchKernelGGL(k_y_ext2int_t, dim3((unsigned)((je - jb + kYT - 1) / kYT)), dim3(256), shm, s, ext, in,
                       p.iperm, p.sframe, p.d_foff, p.C, batch, jb, je, only_frame);
    return;
  }
  hipLaunchKernelGGL(k_y_ext2int, dim3(2048), dim3(256), 0, s, ext, in, p.perm, p.sframe, p.d_foff, p.C, batch,
                     only_frame, p.Mtot);
}
void launch_y_int2ext(const Plan& p, const float2* in, float2* ext, int batch, int only_frame, hipStream_t s) {
  if (p.iperm && batch <= p.C && p.C < 128) {
    int64_t jb, je;
    y_range(p, only_frame, &jb, &je);
    if (je <= jb) return;
    const size_t shm = sizeof(float2) * kYT * (size_t)(p.C + 1);
    hipLaunchKernelGGL(k_y_int2ext_t, dim3((unsigned)((je - jb + kYT - 1) / kYT)), dim3(256), shm, s, in, ext,
                       p.iperm, p.sframe, p.d_foff, p.C, batch, jb, je, only_frame);
    return;
  }
  hipLaunchKernelGGL(k_y_int2ext, dim3(2048), dim3(256), 0, s, in, ext, p.perm, p.sframe, p.d_foff, p.C, batch,
                     only_frame, p.Mtot);
}

// (N^d, batch) column-major <-> X[(f*Ntot + i)*C + c]
__global__ void k_img_ext2coil(const float2* ext, float2* X, int C, int batch, int64_t Ntot, int f) {
  const int64_t tot = Ntot * C;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < tot; idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = idx / C;
    const int c = (int)(idx - i * C);
    X[(int64_t)f * Ntot * C + idx] = (c < batch) ? ext[i + Ntot * c] : make_float2(0.f, 0.f);
  }
}
__global__ void k_img_coil2ext(const float2* X, float2* ext, int C, int batch, int64_t Ntot, int f) {
  const int64_t tot = Ntot * C;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < tot; idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = idx / C;
    const int c = (int)(idx - i * C);
    if (c < batch) ext[i + Ntot * c] = X[(int64_t)f * Ntot * C + idx];
  }
}
// the same as transposes through LDS: 64 pixels per block, each coil's run of
// pixels and each pixel's row of coils contiguous
__global__ void __launch_bounds__(256) k_img_ext2coil_t(const float2* ext, float2* X, int C, int batch,
                                                        int64_t Ntot, int f) {
  extern __shared__ float2 it_[];                // [kYT][C + 1]
  const int64_t i0 = (int64_t)blockIdx.x * kYT;
  const int ni = (int)min<int64_t>(kYT, Ntot - i0), ld = C + 1;
  for (int e = threadIdx.x; e < kYT * batch; e += blockDim.x) {
    const int c = e / kYT, ii = e - c * kYT;
    if (ii < ni) it_[ii * ld + c] = ext[i0 + ii + Ntot * c];
  }
  __syncthreads();
  float2* Xf = X + ((int64_t)f * Ntot + i0) * C;
  for (int e = threadIdx.x; e < kYT * C; e += blockDim.x) {
    const int ii = e / C, c = e - ii * C;
    if (ii < ni) Xf[(int64_t)ii * C + c] = c < batch ? it_[ii * ld + c] : make_float2(0.f, 0.f);
  }
}
__global__ void __launch_bounds__(256) k_img_coil2ext_t(const float2* X, float2* ext, int C, int batch,
                                                        int64_t Ntot, int f) {
  extern __shared__ float2 it_[];                // [kYT][C + 1]
  const int64_t i0 = (int64_t)blockIdx.x * kYT;
  const int ni = (int)min<int64_t>(kYT, Ntot - i0), ld = C + 1;
  const float2* Xf = X + ((int64_t)f * Ntot + i0) * C;
  for (int e = threadIdx.x; e < kYT * batch; e += blockDim.x) {
    const int ii = e / batch, c = e - ii * batch;
    if (ii < ni) it_[ii * ld + c] = Xf[(int64_t)ii * C + c];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < kYT * batch; e += blockDim.x) {
    const int c = e / kYT, ii = e - c * kYT;
    if (ii < ni) ext[i0 + ii + Ntot * c] = it_[ii * ld + c];
  }
}
void launch_img_ext2coil(const Plan& p, const float2* ext, float2* X, int batch, int frame, hipStream_t s) {
  if (p.C < 128) {
    const size_t shm = sizeof(float2) * kYT * (size_t)(p.C + 1);
    hipLaunchKernelGGL(k_img_ext2coil_t, dim3((unsigned)((p.Ntot + kYT - 1) / kYT)), dim3(256), shm, s, ext, X, p.C,
                       batch, p.Ntot, frame);
    return;
  }
  hipLaunchKernelGGL(k_img_ext2coil, dim3(2048), dim3(256), 0, s, ext, X, p.C, batch, p.Ntot, frame);
}
void launch_img_coil2ext(const Plan& p, const float2* X, float2* ext, int batch, int frame, hipStream_t s) {
  if (p.C < 128) {
    const size_t shm = sizeof(float2) * kYT * (size_t)(p.C + 1);
    hipLaunchKernelGGL(k_img_coil2ext_t, dim3((unsigned)((p.Ntot + kYT - 1) / kYT)), dim3(256), shm, s, X, ext, p.C,
                       batch, p.Ntot, frame);
    return;
  }
  hipLaunchKernelGGL(k_img_coil2ext, dim3(2048), dim3(256), 0, s, X, ext, p.C, batch, p.Ntot, frame);
}

// ---------------------------------------------------------------------------
// CGNR (NFFT3 solver with PRECOMPUTE_WEIGHT | PRECOMPUTE_DAMP, NFFTTraits.hpp:122-138)
// batched over (frame, coil) with one set of scalars per pair:
//   cnr[k*F*C + f*C + c], k: 0 dot_r, 1 dot_z, 2 dot_v, 3 alpha, 4 beta, 5 active, 6 new value
constexpr int kCnrBlocks = 256;

// inside NFFT3's damping ball (PRECOMPUTE_DAMP): pixel i of a frame (i < 2^32),
// |(i0, i1[, i2]) - nh|^2 <= nh^2, with 32-bit index arithmetic
__device__ __forceinline__ bool damp_in(uint32_t i, int N0, int N1, int d, int nh) {
  const uint32_t rest = i / (uint32_t)N0;
  const int i0 = (int)(i - rest * (uint32_t)N0);
  const uint32_t r2 = rest / (uint32_t)N1;
  const int i1 = (int)(rest - r2 * (uint32_t)N1);
  int64_t q = (int64_t)(i0 - nh) * (i0 - nh) + (int64_t)(i1 - nh) * (i1 - nh);
  if (d == 3) q += (int64_t)((int)r2 - nh) * ((int)r2 - nh);
  return q <= (int64_t)nh * nh;
}

// per-coil weighted sums over rows of a [rows][C] array; rows are samples of
// each frame (weights w[s]) or pixels of each frame (damping mask)
template <int CP>
__global__ void __launch_bounds__(256) k_cnr_partials(const float2* v, const float* w, const int64_t* foff,
                                                      int64_t Ntot, int mask_nh, int d, int N0, int N1,
                                                      int C, int image, double* partials, const int* fmask) {
  const int f = blockIdx.y;
  if (fmask && fmask[f]) return;
  constexpr int RPB = 256 / CP;
  const int g = threadIdx.x / CP, cl = threadIdx.x % CP;
  const int64_t r0 = image ? (int64_t)f * Ntot : foff[f];
  const int64_t r1 = image ? r0 + Ntot : foff[f + 1];
  __shared__ double sh[256];
  auto weight = [&](int64_t r) -> float {
    if (!image) return w[r];
    return damp_in((uint32_t)(r - r0), N0, N1, d, mask_nh) ? 1.f : 0.f;
  };
  const int64_t stride = (int64_t)gridDim.x * RPB;
  for (int c0 = 0; c0 < C; c0 += CP) {
    const int c = c0 + cl;
    double acc = 0.0;
    if (c < C) {
      // rows r, r + stride, ... summed in that order, eight rows' loads in flight
      constexpr int U = 8;
      int64_t r = r0 + (int64_t)blockIdx.x * RPB + g;
      for (; r + (U - 1) * stride < r1; r += U * stride) {
        float2 a[U];
        float wt[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          a[u] = v[(r + u * stride) * C + c];
          wt[u] = weight(r + u * stride);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) acc += (double)wt[u] * ((double)a[u].x * a[u].x + (double)a[u].y * a[u].y);
      }
      for (; r < r1; r += stride) {
        const float2 a = v[r * C + c];
        acc += (double)weight(r) * ((double)a.x * a.x + (double)a.y * a.y);
      }
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x < CP) {
      double t = 0.0;
      for (int k = 0; k < RPB; ++k) t += sh[k * CP + threadIdx.x];
      if (c0 + threadIdx.x < C) partials[((int64_t)f * gridDim.x + blockIdx.x) * C + c0 + threadIdx.x] = t;
    }
    __syncthreads();
  }
}

// r -= alpha v over the frame's samples, fused with the next dot_r partials:
// the rows, lanes and summation order of k_cnr_partials over the updated r
// (same block partials), so the pass over r is read once
template <int CP>
__global__ void __launch_bounds__(256) k_cnr_update_samples_dot(const float2* rin, float2* r, const float2* v,
                                                                const float* w, const int64_t* foff,
                                                                const double* cnr, int F, int C, double* partials,
                                                                const int* fmask) {
  const int f = blockIdx.y;
  if (fmask && fmask[f]) return;
  constexpr int RPB = 256 / CP, U = 8;
  const int g = threadIdx.x / CP, cl = threadIdx.x % CP;
  const int64_t r0 = foff[f], r1 = foff[f + 1];
  const int64_t stride = (int64_t)gridDim.x * RPB;
  __shared__ double sh[256];
  for (int c0 = 0; c0 < C; c0 += CP) {
    const int c = c0 + cl;
    double acc = 0.0;
    if (c < C) {
      const float al = (float)cnr[(int64_t)3 * F * C + (int64_t)f * C + c];
      int64_t row = r0 + (int64_t)blockIdx.x * RPB + g;
      for (; row + (U - 1) * stride < r1; row += U * stride) {
        float2 a[U], b[U];
        float wt[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          a[u] = rin[(row + u * stride) * C + c];
          b[u] = v[(row + u * stride) * C + c];
          wt[u] = w[row + u * stride];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          a[u].x -= al * b[u].x;
          a[u].y -= al * b[u].y;
          r[(row + u * stride) * C + c] = a[u];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) acc += (double)wt[u] * ((double)a[u].x * a[u].x + (double)a[u].y * a[u].y);
      }
      for (; row < r1; row += stride) {
        float2 a = rin[row * C + c];
        const float2 b = v[row * C + c];
        a.x -= al * b.x;
        a.y -= al * b.y;
        r[row * C + c] = a;
        acc += (double)w[row] * ((double)a.x * a.x + (double)a.y * a.y);
      }
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x < CP) {
      double t = 0.0;
      for (int k = 0; k < RPB; ++k) t += sh[k * CP + threadIdx.x];
      if (c0 + threadIdx.x < C) partials[((int64_t)f * gridDim.x + blockIdx.x) * C + c0 + threadIdx.x] = t;
    }
    __syncthreads();
  }
}

// one block per frame: thread (coil cl, chunk j) sums partials b = j, j + 8, ...
// in order, then chunk sums j = 0..7 are added in order (a fixed order: the
// result does not depend on timing)
// step 1 / 2 (else -1): the scalar step that reads this sum, taken here for the
// same (frame, coil) -- alpha after dot_v, beta after the new dot_z (as k_cnr_scalars)
__global__ void __launch_bounds__(256) k_cnr_reduce(const double* partials, int nb, int C, int F, double* cnr,
                                                    int slot, const int* fmask, int step) {
  constexpr int CB = 32, NJ = 256 / CB;
  __shared__ double sh[256];
  const int f = blockIdx.x;
  if (fmask && fmask[f]) return;
  const int cl = threadIdx.x % CB, j = threadIdx.x / CB;
  for (int c0 = 0; c0 < C; c0 += CB) {
    const int c = c0 + cl;
    double t = 0.0;
    if (c < C) {
#pragma unroll 8
      for (int b = j; b < nb; b += NJ) t += partials[((int64_t)f * nb + b) * C + c];
    }
    sh[threadIdx.x] = t;
    __syncthreads();
    if (j == 0 && c < C) {
      double u = 0.0;
#pragma unroll
      for (int jj = 0; jj < NJ; ++jj) u += sh[jj * CB + cl];
      const int64_t FC = (int64_t)F * C, idx = (int64_t)f * C + c;
      cnr[slot * FC + idx] = u;
      if (step == 1) {          // alpha = dot_z / dot_v (u = dot_v)
        cnr[3 * FC + idx] = (cnr[5 * FC + idx] != 0.0) ? cnr[1 * FC + idx] / u : 0.0;
      } else if (step == 2) {   // beta = new dot_z / dot_z; dot_z = new (u = new dot_z)
        if (cnr[5 * FC + idx] != 0.0) {
          cnr[4 * FC + idx] = u / cnr[1 * FC + idx];
          cnr[1 * FC + idx] = u;
        }
      }
    }
    __syncthreads();
  }
}

static void cnr_dot(const Plan& p, const float2* v, int image, double* dst_cnr, int slot, int step, hipStream_t s) {
  const dim3 g(kCnrBlocks, p.F);
  const int nh = (int)(0.5 * p.N[0]);
  if (p.C <= 8)
    hipLaunchKernelGGL(k_cnr_partials<8>, g, dim3(256), 0, s, v, p.w, p.d_foff, p.Ntot, nh, p.d, p.N[0], p.N[1], p.C, image, p.partials, p.fmask);
  else if (p.C <= 16)
    hipLaunchKernelGGL(k_cnr_partials<16>, g, dim3(256), 0, s, v, p.w, p.d_foff, p.Ntot, nh, p.d, p.N[0], p.N[1], p.C, image, p.partials, p.fmask);
  else
    hipLaunchKernelGGL(k_cnr_partials<32>, g, dim3(256), 0, s, v, p.w, p.d_foff, p.Ntot, nh, p.d, p.N[0], p.N[1], p.C, image, p.partials, p.fmask);
  hipLaunchKernelGGL(k_cnr_reduce, dim3((unsigned)p.F), dim3(256), 0, s, p.partials, kCnrBlocks, p.C, p.F, dst_cnr,
                     slot, p.fmask, step);
}

void launch_cnr_dot_samples(const Plan& p, const float2* v, double* cnr_slot_base, hipStream_t s, int step) {
  // cnr_slot_base encodes the slot: pointer p.cnr + slot*F*C
  const int slot = (int)((cnr_slot_base - p.cnr) / (p.F * p.C));
  cnr_dot(p, v, 0, p.cnr, slot, step, s);
}
void launch_cnr_dot_image(const Plan& p, const float2* v, bool masked, double* cnr_slot_base, hipStream_t s,
                          int step) {
  (void)masked;
  const int slot = (int)((cnr_slot_base - p.cnr) / (p.F * p.C));
  cnr_dot(p, v, 1, p.cnr, slot, step, s);
}

// step 0: active &= dot_r >= eps (NFFTTraits.hpp:131); step 1: alpha; step 2: beta,
// step 3: initialise (active = 1).
__global__ void k_cnr_scalars(double* cnr, int FC, int step, double eps, const int* fmask, int C) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= FC) return;
  if (fmask && fmask[idx / C]) return;
  double* dot_r = cnr + 0 * FC;
  double* dot_z = cnr + 1 * FC;
  double* dot_v = cnr + 2 * FC;
  double* alpha = cnr + 3 * FC;
  double* beta = cnr + 4 * FC;
  double* active = cnr + 5 * FC;
  double* nv = cnr + 6 * FC;
  if (step == 3) {
    active[idx] = 1.0;
    alpha[idx] = 0.0;
    beta[idx] = 0.0;
  } else if (step == 0) {
    if (!(dot_r[idx] >= eps)) active[idx] = 0.0;
  } else if (step == 1) {
    alpha[idx] = (active[idx] != 0.0) ? dot_z[idx] / dot_v[idx] : 0.0;
  } else {
    if (active[idx] != 0.0) {
      beta[idx] = nv[idx] / dot_z[idx];
      dot_z[idx] = nv[idx];
    }
  }
}

void launch_cnr_scalars(const Plan& p, int step, double eps, hipStream_t s) {
  const int n = p.F * p.C;
  hipLaunchKernelGGL(k_cnr_scalars, dim3((n + 255) / 256), dim3(256), 0, s, p.cnr, n, step, eps, p.fmask, p.C);
}

// step 0: f_hat += alpha w_hat p_hat ; step 1: p_hat = active ? beta p_hat + z : p_hat.
// Blocks (pixel range, frame), lanes (pixel, coil): no index division per element.
template <int CP>
__global__ void __launch_bounds__(256) k_cnr_update_image(float2* fhat, float2* phat, const float2* z,
                                                          const double* cnr, int F, int C, int64_t Ntot, int mask_nh,
                                                          int d, int N0, int N1, int step, const int* fmask) {
  const int f = blockIdx.y;
  if (fmask && fmask[f]) return;
  constexpr int RPB = 256 / CP;
  const int g = threadIdx.x / CP, cl = threadIdx.x % CP;
  const int64_t stride = (int64_t)gridDim.x * RPB;
  const int64_t FC = (int64_t)F * C;
  for (int c0 = 0; c0 < C; c0 += CP) {
    const int c = c0 + cl;
    if (c >= C) continue;
    const int64_t fc = (int64_t)f * C + c;
    float2* fh = fhat + (int64_t)f * Ntot * C + c;
    float2* ph = phat + (int64_t)f * Ntot * C + c;
    if (step == 0) {
      const float al = (float)cnr[3 * FC + fc];
      for (int64_t i = (int64_t)blockIdx.x * RPB + g; i < Ntot; i += stride) {
        if (!damp_in((uint32_t)i, N0, N1, d, mask_nh)) continue;
        const float2 pp = ph[i * C];
        float2 ff = fh[i * C];
        ff.x += al * pp.x;
        ff.y += al * pp.y;
        fh[i * C] = ff;
      }
    } else {
      if (cnr[5 * FC + fc] == 0.0) continue;
      const float be = (float)cnr[4 * FC + fc];
      const float2* zz = z + (int64_t)f * Ntot * C + c;
      for (int64_t i = (int64_t)blockIdx.x * RPB + g; i < Ntot; i += stride) {
        const float2 zv = zz[i * C];
        float2 pp = ph[i * C];
        pp.x = be * pp.x + zv.x;
        pp.y = be * pp.y + zv.y;
        ph[i * C] = pp;
      }
    }
  }
}

void launch_cnr_update_image(const Plan& p, float2* fhat, float2* phat, const float2* z, int step, hipStream_t s) {
  const dim3 g((unsigned)std::max(64, 2048 / std::max(1, p.F)), (unsigned)p.F);
  const int nh = (int)(0.5 * p.N[0]);
  if (p.C <= 8)
    hipLaunchKernelGGL(k_cnr_update_image<8>, g, dim3(256), 0, s, fhat, phat, z, p.cnr, p.F, p.C, p.Ntot, nh, p.d,
                       p.N[0], p.N[1], step, p.fmask);
  else if (p.C <= 16)
    hipLaunchKernelGGL(k_cnr_update_image<16>, g, dim3(256), 0, s, fhat, phat, z, p.cnr, p.F, p.C, p.Ntot, nh, p.d,
                       p.N[0], p.N[1], step, p.fmask);
  else
    hipLaunchKernelGGL(k_cnr_update_image<32>, g, dim3(256), 0, s, fhat, phat, z, p.cnr, p.F, p.C, p.Ntot, nh, p.d,
                       p.N[0], p.N[1], step, p.fmask);
}

// r = rin - alpha v (rin == r: in place), then dot_r (slot 0) from the fused pass's partials
void launch_cnr_update_samples_dot(const Plan& p, const float2* rin, float2* r, const float2* v, hipStream_t s) {
  const dim3 g(kCnrBlocks, p.F);
  if (p.C <= 8)
    hipLaunchKernelGGL(k_cnr_update_samples_dot<8>, g, dim3(256), 0, s, rin, r, v, p.w, p.d_foff, p.cnr, p.F, p.C,
                       p.partials, p.fmask);
  else if (p.C <= 16)
    hipLaunchKernelGGL(k_cnr_update_samples_dot<16>, g, dim3(256), 0, s, rin, r, v, p.w, p.d_foff, p.cnr, p.F, p.C,
                       p.partials, p.fmask);
  else
    hipLaunchKernelGGL(k_cnr_update_samples_dot<32>, g, dim3(256), 0, s, rin, r, v, p.w, p.d_foff, p.cnr, p.F, p.C,
                       p.partials, p.fmask);
  hipLaunchKernelGGL(k_cnr_reduce, dim3((unsigned)p.F), dim3(256), 0, s, p.partials, kCnrBlocks, p.C, p.F, p.cnr, 0,
                     p.fmask, -1);
}

// Toeplitz spectrum: grid order Khat[f][r + cells*pos] -> conv-pass layout
// That[(f*cells + r)*nl + pos] = Re(Khat) * scale
__global__ void k_toep_layout(const float2* Khat, float* That, int64_t cells, int nl, int F, float scale) {
  const int64_t tot = (int64_t)F * cells * nl;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < tot; idx += (int64_t)gridDim.x * blockDim.x) {
    const int pos = (int)(idx % nl);
    const int64_t fr = idx / nl;
    const int64_t r = fr % cells, f = fr / cells;
    That[idx] = Khat[f * cells * nl + r + cells * pos].x * scale;
  }
}

void launch_toep_layout(const float2* Khat, float* That, int64_t cells, int nl, int F, float scale, hipStream_t s) {
  hipLaunchKernelGGL(k_toep_layout, dim3(2048), dim3(256), 0, s, Khat, That, cells, nl, F, scale);
}

void launch_copy(void* dst, const void* src, size_t bytes, hipStream_t s) {
  (void)hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s);
}

__global__ void k_fill_int(int* dst, int v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = v;
}
void launch_fill_int(int* dst, int value, int n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_fill_int, dim3((n + 255) / 256), dim3(256), 0, s, dst, value, n);
}

// zero-fill of n 32-bit words (a kernel, not a memset node: safe where a
// caller captures the stream into a graph)
__global__ void k_zero_u32(uint32_t* dst, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = 0u;
}
void launch_zero(void* dst, size_t bytes, hipStream_t s) {
  const int64_t n = (int64_t)(bytes / 4);
  if (n <= 0) return;
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)blocks), dim3(256), 0, s, (uint32_t*)dst, n);
}

// The CG tail of the last iteration when the loop ends with it still pending
// (the folded tail runs in the next pad pass, and there is none): per frame not
// yet stopped, alpha and ||r_new||^2 from the crop's per-row partials in the
// pad's order, x += alpha p, r -= alpha q (k_cg_fold_update), then the frame's
// scalars and stop state (k_cg_fold_commit, one block per frame).
__device__ __forceinline__ void fold_scalars(const double* fp, const double* fr, const double* fq, const double* frr,
                                             int nrpart, double* al, double* g, double* R) {
  // lanes 0..63 of the calling wave; the result in every lane
  double sp = 0.0, sr = 0.0, sq = 0.0, sr2 = 0.0;
  for (int u = (int)(threadIdx.x & 63); u < nrpart; u += 64) {
    sp += fp[u];
    sr += fr[u];
    sq += fq[u];
    sr2 += frr[u];
  }
  sp = wave_sum_d(sp);
  sr = wave_sum_d(sr);
  sq = wave_sum_d(sq);
  sr2 = wave_sum_d(sr2);
  *R = sr2;
  *al = sr2 / sp;
  *g = fmax(sr2 - 2.0 * (*al) * sr + (*al) * (*al) * sq, 0.0);
}

__global__ void __launch_bounds__(256) k_cg_fold_update(float2* x, float2* r, const float2* pv, const float2* qv,
                                                        const CgState* st, const double* fpq, const double* frq,
                                                        const double* fqq, const double* frr, int nrpart,
                                                        int64_t Ntot) {
  const int f = blockIdx.y;
  const CgState& cs = st[f];
  if (cs.done || !cs.bpend) return;
  __shared__ float al_s;
  if (threadIdx.x < 64) {
    double al, g, R;
    const int64_t o = (int64_t)f * nrpart;
    fold_scalars(fpq + o, frq + o, fqq + o, frr + o, nrpart, &al, &g, &R);
    if (threadIdx.x == 0) al_s = (float)al;
  }
  __syncthreads();
  const float al = al_s;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < Ntot; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t o = (int64_t)f * Ntot + i;
    const float2 pp = pv[o], qq = qv[o];
    float2 xx = x[o], rr = r[o];
    xx.x += al * pp.x; xx.y += al * pp.y;
    rr.x -= al * qq.x; rr.y -= al * qq.y;
    x[o] = xx;
    r[o] = rr;
  }
}

__global__ void k_cg_fold_commit(CgState* st, const double* fpq, const double* frq, const double* fqq,
                                 const double* frr, int nrpart, int tolkind, double tol, int maxit, double* hist,
                                 int hcap) {
  const int f = blockIdx.x;
  CgState& cs = st[f];
  if (cs.done || !cs.bpend) return;
  double al, g, R;
  const int64_t o = (int64_t)f * nrpart;
  fold_scalars(fpq + o, frq + o, fqq + o, frr + o, nrpart, &al, &g, &R);
  if (threadIdx.x == 0) {
    const int it = cs.it_snap + 1;
    const double rel = (tolkind == CGS_TOL_SQUARED_RATIO) ? g / cs.xn : sqrt(g / cs.bn);
    const int nan = (cs.nan || isnan(rel) || isnan(al)) ? 1 : 0;
    cs.alpha = al;
    cs.beta = g / R;
    cs.rn = g;
    cs.it = it;
    cs.rel = rel;
    cs.nan = nan;
    cs.done = (nan || rel <= tol || it >= maxit) ? 1 : 0;
    cs.bpend = 0;
    if (hist && it < hcap) hist[(int64_t)f * hcap + it] = rel;
  }
}

void launch_cg_fold_final(const Plan& p, float2* x, float2* r, const float2* pv, const float2* qv, const double* fpq,
                          const double* frq, const double* fqq, const double* frr, int nrpart, int tolkind,
                          double tol, int maxit, hipStream_t s) {
  const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((p.Ntot + 255) / 256, 1024));
  hipLaunchKernelGGL(k_cg_fold_update, dim3((unsigned)blocks, (unsigned)p.F), dim3(256), 0, s, x, r, pv, qv, p.cgst,
                     fpq, frq, fqq, frr, nrpart, p.Ntot);
  hipLaunchKernelGGL(k_cg_fold_commit, dim3((unsigned)p.F), dim3(64), 0, s, p.cgst, fpq, frq, fqq, frr, nrpart,
                     tolkind, tol, maxit, p.hist, p.hist_cap);
}

// Maps already on the device, (Ntot, nm, C) column-major -> [z][pixel][coil],
// and the intensity map 1 / (sqrt(sum_c |S_c|^2) + 1e-9) accumulated in FP32 in
// coil order with separately rounded products and sums, as the host loader does.
__global__ void k_maps_load(const float2* __restrict__ S, float2* __restrict__ Sd, float* __restrict__ I,
                            int64_t Ntot, int64_t nm, int C) {
  const int64_t n = Ntot * nm;
  for (int64_t zi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; zi < n; zi += (int64_t)gridDim.x * blockDim.x) {
    const int64_t z = zi / Ntot, i = zi - z * Ntot;
    float acc = 0.f;
    for (int c = 0; c < C; ++c) {
      const float2 v = S[i + Ntot * (z + nm * c)];
      Sd[zi * C + c] = v;
      acc = __fadd_rn(acc, __fadd_rn(__fmul_rn(v.x, v.x), __fmul_rn(v.y, v.y)));
    }
    I[zi] = (float)(1.0 / (sqrt((double)acc) + 1.e-9));
  }
}
void launch_maps_load(const float2* S, float2* Sd, float* I, int64_t Ntot, int64_t nm, int C, hipStream_t s) {
  const int64_t n = Ntot * nm;
  const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192));
  hipLaunchKernelGGL(k_maps_load, dim3((unsigned)blocks), dim3(256), 0, s, S, Sd, I, Ntot, nm, C);
}

}  // namespace cgs
