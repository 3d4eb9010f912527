// kernels_bin.hip -- binned gridding: the adjoint window scatter (B^T of
// nfft_adjoint, NFFTTraits.hpp:201-204) with every sample read once.
//
// Samples are grouped by the bin (TB^d grid cells) that holds their base cell
// (the first cell of their footprint).  A bin's region is the bin plus the
// taps - 1 cells above it on every axis (E = TB + taps - 1 cells per axis): it
// holds the whole footprint of every sample of the bin.  A work item (the
// samples of one bin, or a chunk of a heavy bin's) accumulates its samples'
// footprints, for CG coils, into the region in LDS and writes the region to its
// own slot of a region buffer.  k_grid_bin_sum then forms every grid cell as the
// sum of the region cells that map onto it -- its own bin's, and those of the
// bin below it on each axis whose region reaches it -- over the items of each
// bin, in a fixed order.  No zero-fill of the grid and no global atomics.
//
// In LDS every (coil, re/im) has a plane of E^d floats with rows padded to P1
// floats (P1 % 32 = 8 or 24, so that the 64 lanes of a footprint row-plane fall
// on every LDS bank exactly twice).  The lanes of a wave are the footprint taps
// (s0, s1): a wave takes one sample at a time, forms its lanes' window weights
// once and adds weight x datum for each coil (and, in 3D, each tap s2 along the
// third axis) with LDS float atomics, since the waves of the block apply
// different samples at the same time.  The order of the LDS additions is not
// fixed: results vary in the last bits from run to run (the plan's
// CGS_DETERMINISTIC flag keeps the fixed-order kernels).
#include <algorithm>
#include <cstdlib>

#include "cgs_internal.hpp"

namespace cgs {

// smallest row stride >= E whose residue mod 32 is 8 or 24
int bin_row_stride(int E) {
  int P = E;
  while ((P & 31) != 8 && (P & 31) != 24) ++P;
  return P;
}

static int64_t bin_plane(int d, int E, int P1) {
  return (int64_t)P1 * E * (d == 3 ? E : 1);
}

// samples staged per batch: 2D 128, 3D 256
static int bin_batch(int d) { return d == 3 ? 256 : 128; }

size_t grid_bin_lds(int d, int taps, int CG, int E, int P1) {
  const int B = bin_batch(d);
  return sizeof(float) * (2 * (size_t)CG * (size_t)bin_plane(d, E, P1) + (size_t)B * (1 + d * taps) +
                          (size_t)B * 2 * CG);
}

// Work item = (frame, bin, ids [first, end)), blockIdx.y = coil group.  The
// block stages B samples at a time in LDS (records from brec, data and weight
// through the ids), the next batch's loads in flight while the waves apply the
// current one; a wave applies one sample at a time.
template <int CG, int TAPS, int D, int NT>
__global__ void __launch_bounds__(NT) k_grid_bin(BinArgs a) {
  constexpr int RS = 64 / TAPS;                 // lane = s0 + TAPS * rs
  constexpr int NPASS = (TAPS + RS - 1) / RS;   // passes over the footprint rows s1
  constexpr int NW = NT / 64;
  constexpr int B = (D == 3) ? 256 : 128;
  constexpr int R = 1 + D * TAPS;               // record words per sample
  constexpr int UR = (B * R + NT - 1) / NT;     // record words per thread and batch
  constexpr int UY = (B * CG + NT - 1) / NT;    // data values per thread and batch
  extern __shared__ float lds[];
  const int4 it = a.items[blockIdx.x];          // {frame * nbins + bin, first id, end id, slot}
  const int f = (int)(it.x / a.nbins);
  if (a.fdone && a.fdone[f]) return;            // block-uniform
  const int g = blockIdx.y, cbase = g * CG;
  const int E = a.E, P1 = a.P1, P2 = P1 * E;
  const int PC = (D == 3) ? P2 * E : P2;        // floats of one (coil, component) plane
  float* srec = lds + 2 * CG * PC;              // [B][R]
  float2* sy = (float2*)(srec + B * R);         // [B][CG]
  for (int i = threadIdx.x; i < 2 * CG * PC; i += NT) lds[i] = 0.f;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int s0 = lane % TAPS, rs = lane / TAPS;
  const int beg = it.y, end = it.z;
  const int nbat = (end - beg + B - 1) / B;
  float rv[UR];
  float2 yv[UY];
  auto load = [&](int bi) {
    const int k0 = beg + bi * B, nb = min(B, end - k0);
#pragma unroll
    for (int u = 0; u < UR; ++u) {
      const int idx = threadIdx.x + NT * u;
      rv[u] = (idx < nb * R) ? a.brec[(int64_t)k0 * R + idx] : 0.f;
    }
#pragma unroll
    for (int u = 0; u < UY; ++u) {
      const int idx = threadIdx.x + NT * u;
      yv[u] = make_float2(0.f, 0.f);
      if (idx < nb * CG) {
        const int j = idx / CG, c = idx - j * CG;
        if (cbase + c < a.C) {
          const int s = a.ids[k0 + j];
          yv[u] = a.y[(int64_t)s * a.C + cbase + c];
          if (a.w) { const float ww = a.w[s]; yv[u].x *= ww; yv[u].y *= ww; }
        }
      }
    }
  };
  if (nbat > 0) load(0);
  for (int bi = 0; bi < nbat; ++bi) {
    const int nb = min(B, end - beg - bi * B);
    __syncthreads();                            // the previous batch is consumed (and the region zeroed)
#pragma unroll
    for (int u = 0; u < UR; ++u) {
      const int idx = threadIdx.x + NT * u;
      if (idx < B * R) srec[idx] = rv[u];
    }
#pragma unroll
    for (int u = 0; u < UY; ++u) {
      const int idx = threadIdx.x + NT * u;
      if (idx < B * CG) sy[idx] = yv[u];
    }
    __syncthreads();
    if (bi + 1 < nbat) load(bi + 1);            // in flight while this batch is applied
    for (int j = wave; j < nb; j += NW) {
      const float* rec = srec + j * R;
      const int off = __float_as_int(rec[0]);
      const int r0 = off & 0xff, r1 = (off >> 8) & 0xff, r2 = (off >> 16) & 0xff;
      float2 y[CG];
#pragma unroll
      for (int c = 0; c < CG; ++c) y[c] = sy[j * CG + c];
      const float px = rec[1 + s0];
#pragma unroll
      for (int pp = 0; pp < NPASS; ++pp) {
        const int s1 = rs + pp * RS;
        if (rs < RS && s1 < TAPS) {
          const float pxy = px * rec[1 + TAPS + s1];
          const int l = (r0 + s0) + P1 * (r1 + s1);
          if (D == 2) {
#pragma unroll
            for (int c = 0; c < CG; ++c) {
              atomicAdd(&lds[(2 * c) * PC + l], pxy * y[c].x);
              atomicAdd(&lds[(2 * c + 1) * PC + l], pxy * y[c].y);
            }
          } else {
#pragma unroll
            for (int s2 = 0; s2 < TAPS; ++s2) {
              const float w3 = pxy * rec[1 + 2 * TAPS + s2];
              const int l3 = l + P2 * (r2 + s2);
#pragma unroll
              for (int c = 0; c < CG; ++c) {
                atomicAdd(&lds[(2 * c) * PC + l3], w3 * y[c].x);
                atomicAdd(&lds[(2 * c + 1) * PC + l3], w3 * y[c].y);
              }
            }
          }
        }
      }
    }
  }
  __syncthreads();
  // the region, all of it, to the item's slot: [slot][group][E^d][CG]
  float2* out = a.reg + ((int64_t)it.w * a.ngroups + g) * a.ecells * CG;
  for (int64_t idx = threadIdx.x; idx < a.ecells * CG; idx += NT) {
    const int c = (int)(idx % CG);
    const int64_t e = idx / CG;
    const int e0 = (int)(e % E);
    const int64_t er = e / E;
    const int e1 = (int)(er % E), e2 = (D == 3) ? (int)(er / E) : 0;
    const int l = e0 + P1 * e1 + P2 * e2;
    out[idx] = make_float2(lds[(2 * c) * PC + l], lds[(2 * c + 1) * PC + l]);
  }
}

// Every grid cell of every active frame: the sum over the (bin, region cell)
// pairs that map onto it -- per axis the cell's own bin at offset g mod TB, and
// the bin below at offset TB + g mod TB while that is inside the region -- and
// over each bin's items, in that order.  Blocks take NCB consecutive cells and
// all coils, staged in LDS so that the grid rows are written whole.
template <int D>
__global__ void __launch_bounds__(256) k_grid_bin_sum(BinArgs a, int ncb) {
  extern __shared__ float2 tile[];   // [ncb][C]
  const int64_t chunks = (a.ntot + ncb - 1) / ncb;
  const int f = (int)(blockIdx.x / chunks);
  const int64_t cell0 = (int64_t)(blockIdx.x % chunks) * ncb;
  if (a.fdone && a.fdone[f]) return;
  const int ncell = (int)min<int64_t>(ncb, a.ntot - cell0);
  const int C = a.C, CG = a.CG, E = a.E, TB = a.TB;
  const int tot = a.ngroups * CG * ncell;
  for (int idx = threadIdx.x; idx < tot; idx += blockDim.x) {
    const int cc = idx % CG;
    const int j = (idx / CG) % ncell;
    const int g = idx / (CG * ncell);
    const int c = g * CG + cc;
    if (c >= C) continue;
    const int64_t cell = cell0 + j;
    const int gx[3] = {(int)(cell % a.n0), (int)((cell / a.n0) % a.n1), (D == 3) ? (int)(cell / ((int64_t)a.n0 * a.n1)) : 0};
    const int nb[3] = {a.nb0, a.nb1, a.nb2};
    int tb[3][2], eo[3][2], nc[3] = {1, 1, 1};
#pragma unroll
    for (int ax = 0; ax < D; ++ax) {
      const int t = gx[ax] / TB, e = gx[ax] - t * TB;
      tb[ax][0] = t;
      eo[ax][0] = e;
      if (e + TB < E) {
        tb[ax][1] = (t == 0) ? nb[ax] - 1 : t - 1;
        eo[ax][1] = e + TB;
        nc[ax] = 2;
      }
    }
    if (D == 2) { tb[2][0] = 0; eo[2][0] = 0; }
    float2 acc = make_float2(0.f, 0.f);
    for (int i2 = 0; i2 < nc[2]; ++i2)
      for (int i1 = 0; i1 < nc[1]; ++i1)
        for (int i0 = 0; i0 < nc[0]; ++i0) {
          const int64_t bin = tb[0][i0] + (int64_t)a.nb0 * (tb[1][i1] + (int64_t)a.nb1 * tb[2][i2]);
          const int64_t e = eo[0][i0] + (int64_t)E * (eo[1][i1] + (int64_t)E * eo[2][i2]);
          const int2 r = a.range[(int64_t)f * a.nbins + bin];
          for (int q = 0; q < r.y; ++q) {
            const float2 v = a.reg[(((int64_t)(r.x + q) * a.ngroups + g) * a.ecells + e) * CG + cc];
            acc.x += v.x;
            acc.y += v.y;
          }
        }
    tile[j * C + c] = acc;
  }
  __syncthreads();
  float2* dst = a.grid + ((int64_t)f * a.ntot + cell0) * C;
  for (int idx = threadIdx.x; idx < ncell * C; idx += blockDim.x) dst[idx] = tile[idx];
}

template <int CG, int TAPS, int D, int NT>
static void grid_bin_go(const BinArgs& a, dim3 g, size_t shm, hipStream_t s) {
  // more than 64 KB of dynamic LDS needs the kernel's limit raised (once)
  static const bool raised = [] {
    return hipFuncSetAttribute(reinterpret_cast<const void*>(&k_grid_bin<CG, TAPS, D, NT>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024) == hipSuccess;
  }();
  (void)raised;
  hipLaunchKernelGGL((k_grid_bin<CG, TAPS, D, NT>), g, dim3(NT), shm, s, a);
}

template <int CG, int D, int NT>
static void grid_bin_taps(const BinArgs& a, dim3 g, size_t shm, hipStream_t s) {
  switch (a.taps) {
    case 4: grid_bin_go<CG, 4, D, NT>(a, g, shm, s); break;
    case 6: grid_bin_go<CG, 6, D, NT>(a, g, shm, s); break;
    case 8: grid_bin_go<CG, 8, D, NT>(a, g, shm, s); break;
    case 10: grid_bin_go<CG, 10, D, NT>(a, g, shm, s); break;
    default: break;
  }
}

template <int NT>
static void grid_bin_nt(const BinArgs& a, dim3 g, size_t shm, hipStream_t s) {
  if (a.d == 2) {
    if (a.CG == 8) grid_bin_taps<8, 2, NT>(a, g, shm, s);
    else grid_bin_taps<4, 2, NT>(a, g, shm, s);
  } else {
    if (a.CG == 2) grid_bin_taps<2, 3, NT>(a, g, shm, s);
    else grid_bin_taps<1, 3, NT>(a, g, shm, s);
  }
}

// CG: 2D 4 or 8, 3D 1 or 2; threads 256, 512 or 1024
void launch_grid_bin(const BinArgs& a, int threads, hipStream_t s) {
  if (a.nitems == 0) return;
  const size_t shm = grid_bin_lds(a.d, a.taps, a.CG, a.E, a.P1);
  const dim3 g((unsigned)a.nitems, (unsigned)a.ngroups);
  if (threads >= 1024) grid_bin_nt<1024>(a, g, shm, s);
  else if (threads >= 512) grid_bin_nt<512>(a, g, shm, s);
  else grid_bin_nt<256>(a, g, shm, s);
}

void launch_grid_bin_sum(const BinArgs& a, hipStream_t s) {
  const int ncb = (int)std::max<int64_t>(1, std::min<int64_t>(64, 8192 / std::max(1, a.C)));
  const int64_t chunks = (a.ntot + ncb - 1) / ncb;
  const size_t shm = sizeof(float2) * (size_t)ncb * a.C;
  const dim3 g((unsigned)(chunks * a.F));
  if (a.d == 3) hipLaunchKernelGGL((k_grid_bin_sum<3>), g, dim3(256), shm, s, a, ncb);
  else hipLaunchKernelGGL((k_grid_bin_sum<2>), g, dim3(256), shm, s, a, ncb);
}

}  // namespace cgs
