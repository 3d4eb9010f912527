// kernels_grid.hip -- Kaiser-Bessel window gather (forward B) and tile-binned
// window scatter (adjoint B^T) on the coil-interleaved oversampled grid.
//
// Both kernels put the coils of one sample across the lanes of a wave:
// lane = cl + CG*rs with cl the coil within a group of CG coils and rs one of
// RS = 64/CG tap rows handled side by side.  The window weights of a sample are
// wave-uniform (scalar loads); every tap is one coalesced CG*8-byte access.
//
// Gridding is "owner computes": each grid tile (T^d cells) is owned by one
// work item that visits every sample whose 2m+2 footprint reaches into the
// tile and accumulates only the taps that land inside it, in LDS.  Each cell is
// therefore written exactly once, with no atomics, no zero-fill of the grid and
// a fixed summation order (deterministic).  Tiles whose visit lists are long
// (the k-space centre of radial trajectories) are split across several items
// that write partial tiles, summed afterwards in a fixed order.
#include <cstdlib>

#include "cgs_internal.hpp"

namespace cgs {

typedef float gf32x16 __attribute__((ext_vector_type(16)));   // MFMA 32x32 accumulator

// v + (v of the lane that the DPP control CTRL selects); lanes whose source
// lies outside the 16-lane row add zero
template <int CTRL>
__device__ __forceinline__ float2 dpp_add2(float2 v) {
  const float x = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v.x), CTRL, 0xF, 0xF, true));
  const float y = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v.y), CTRL, 0xF, 0xF, true));
  return make_float2(v.x + x, v.y + y);
}


// ---------------------------------------------------------------------------
// forward B: y[s][c] = sum_taps psi * G[cell][c]
template <int CG, int TAPS, int D>
__global__ void __launch_bounds__(256) k_interp(SampArgs a) {
  constexpr int RS = 64 / CG;
  constexpr int ROWS = (D == 3) ? TAPS * TAPS : TAPS;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int cl = lane % CG, rs = lane / CG;
  const int item = blockIdx.x * 4 + wave;
  if (item >= a.nitems) return;
  const int4 it = a.items[item];
  if (a.fdone && a.fdone[it.x]) return;
  const float2* g = a.grid + (int64_t)it.x * a.ntot * a.C;
  const int n0 = a.n0, n1 = a.n1, n2 = a.n2;
  for (int s = it.y; s < it.y + it.z; ++s) {
    const int* bs = a.base + (int64_t)s * D;
    const float* ps = a.psi + (int64_t)s * D * TAPS;
    const int b0 = bs[0], b1 = bs[1], b2 = (D == 3) ? bs[2] : 0;
    for (int c0 = 0; c0 < a.C; c0 += CG) {
      const int c = c0 + cl;
      float2 acc = make_float2(0.f, 0.f);
      if (c < a.C) {
        for (int r = rs; r < ROWS; r += RS) {
          const int s1 = (D == 3) ? r % TAPS : r;
          int l1 = b1 + s1;
          if (l1 >= n1) l1 -= n1;
          float wr = ps[TAPS + s1];
          int64_t row = l1;
          if (D == 3) {
            const int s2 = r / TAPS;
            int l2 = b2 + s2;
            if (l2 >= n2) l2 -= n2;
            wr *= ps[2 * TAPS + s2];
            row += (int64_t)n1 * l2;
          }
          const float2* gr = g + row * n0 * a.C + c;
          float2 racc = make_float2(0.f, 0.f);
#pragma unroll
          for (int s0 = 0; s0 < TAPS; ++s0) {
            int l0 = b0 + s0;
            if (l0 >= n0) l0 -= n0;
            const float2 v = gr[(int64_t)l0 * a.C];
            const float p0 = ps[s0];
            racc.x += p0 * v.x;
            racc.y += p0 * v.y;
          }
          acc.x += wr * racc.x;
          acc.y += wr * racc.y;
        }
      }
      acc = make_float2(class_sum<CG>(acc.x), class_sum<CG>(acc.y));   // over the row slots
      if (rs == 0 && c < a.C) a.y[(int64_t)s * a.C + c] = acc;
    }
  }
}

// ---------------------------------------------------------------------------
// adjoint B^T, owner-computes tile accumulation.  Block = `waves` waves; wave
// v handles coils [cbase + v*CG, +CG); LDS tile [T^d cells][CGW coils] with a
// padded row stride.  Only one wave touches a given (cell, coil) and a wave
// visits samples in order, so plain read-modify-write in LDS is race-free.
template <int CG, int TAPS, int D>
__global__ void __launch_bounds__(256) k_grid(GridArgs a) {
  constexpr int RS = 64 / CG;
  extern __shared__ float2 tile[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int cl = lane % CG, rs = lane / CG;
  const GridItem it = a.items[blockIdx.x];
  if (a.fdone && a.fdone[it.frame]) return;
  const int T = a.T, rowstride = a.rowstride;
  const int cbase = blockIdx.y * a.CGW;
  const int cw = wave * CG + cl;            // coil slot in the tile
  const int c = cbase + cw;
  const bool cok = (cw < a.CGW) && (c < a.C);

  // tile coordinates
  int t0 = it.tile % a.nt0, t1 = (it.tile / a.nt0) % a.nt1, t2 = (D == 3) ? it.tile / (a.nt0 * a.nt1) : 0;
  const int ts0 = t0 * T, ts1 = t1 * T, ts2 = t2 * T;
  const int te0 = min(T, a.n0 - ts0), te1 = min(T, a.n1 - ts1), te2 = (D == 3) ? min(T, a.n2 - ts2) : 1;
  const int nrows_tile = (D == 3) ? T * T : T;

  // zero the tile
  for (int idx = threadIdx.x; idx < nrows_tile * rowstride; idx += blockDim.x)
    tile[idx] = make_float2(0.f, 0.f);
  lds_barrier();

  for (int v = it.vbeg; v < it.vend; ++v) {
    const int s = a.visits[v];
    const int* bs = a.base + (int64_t)s * D;
    const float* ps = a.psi + (int64_t)s * D * TAPS;
    int rel0 = bs[0] - ts0; if (rel0 < 0) rel0 += a.n0; if (rel0 >= te0) rel0 -= a.n0;
    int rel1 = bs[1] - ts1; if (rel1 < 0) rel1 += a.n1; if (rel1 >= te1) rel1 -= a.n1;
    int rel2 = 0;
    if (D == 3) { rel2 = bs[2] - ts2; if (rel2 < 0) rel2 += a.n2; if (rel2 >= te2) rel2 -= a.n2; }
    const int lo0 = max(0, -rel0), hi0 = min(TAPS, te0 - rel0);
    const int lo1 = max(0, -rel1), hi1 = min(TAPS, te1 - rel1);
    const int lo2 = (D == 3) ? max(0, -rel2) : 0, hi2 = (D == 3) ? min(TAPS, te2 - rel2) : 1;
    const int nr1 = hi1 - lo1;
    const int nrows = nr1 * (hi2 - lo2);
    if (nr1 <= 0 || hi0 <= lo0 || nrows <= 0) continue;
    float2 yv = make_float2(0.f, 0.f);
    if (cok) {
      yv = a.y[(int64_t)s * a.C + c];
      if (a.w) { const float ww = a.w[s]; yv.x *= ww; yv.y *= ww; }
    }
    for (int r = rs; r < nrows; r += RS) {
      const int s1 = lo1 + r % nr1;
      float wr = ps[TAPS + s1];
      int trow = rel1 + s1;
      if (D == 3) {
        const int s2 = lo2 + r / nr1;
        wr *= ps[2 * TAPS + s2];
        trow += (rel2 + s2) * T;
      }
      const float vx = yv.x * wr, vy = yv.y * wr;
      float2* dst = tile + trow * rowstride + cw;
      for (int s0 = lo0; s0 < hi0; ++s0) {
        const float p0 = ps[s0];
        float2* cell = dst + (rel0 + s0) * a.CGW;
        float2 cur = *cell;
        cur.x += p0 * vx;
        cur.y += p0 * vy;
        *cell = cur;
      }
    }
  }
  lds_barrier();

  // write the owned tile (all coils of the group), or a partial slot
  const int ncell = te0 * te1 * te2;
  for (int idx = threadIdx.x; idx < ncell * a.CGW; idx += blockDim.x) {
    const int cc = idx % a.CGW, cell = idx / a.CGW;
    const int cg = cbase + cc;
    if (cg >= a.C) continue;
    const int x0 = cell % te0, x1 = (cell / te0) % te1, x2 = cell / (te0 * te1);
    const float2 v = tile[(x1 + x2 * T) * rowstride + x0 * a.CGW + cc];
    if (it.slot < 0) {
      const int64_t l = (ts0 + x0) + (int64_t)a.n0 * ((ts1 + x1) + (int64_t)a.n1 * (ts2 + x2));
      a.grid[((int64_t)it.frame * a.ntot + l) * a.C + cg] = v;
    } else {
      const int64_t lc = x0 + (int64_t)T * (x1 + (int64_t)T * x2);
      a.partial[((int64_t)it.slot * a.tilecells + lc) * a.C + cg] = v;
    }
  }
}

// sums the partial tiles of the heavy tiles in slot order
__global__ void k_grid_reduce(const ReduceItem* items, const float2* partial, float2* grid,
                              int C, int T, int d, int n0, int n1, int n2, int nt0, int nt1,
                              int64_t ntot, int64_t tilecells, const int* fdone) {
  // blockIdx.y splits the tile's (cell, coil) outputs so heavy tiles (the
  // k-space centre: thousands of slots in 3D) spread over many blocks
  const ReduceItem it = items[blockIdx.x];
  if (fdone && fdone[it.frame]) return;
  const int t0 = it.tile % nt0, t1 = (it.tile / nt0) % nt1, t2 = (d == 3) ? it.tile / (nt0 * nt1) : 0;
  const int ts0 = t0 * T, ts1 = t1 * T, ts2 = t2 * T;
  const int te0 = min(T, n0 - ts0), te1 = min(T, n1 - ts1), te2 = (d == 3) ? min(T, n2 - ts2) : 1;
  const int ncell = te0 * te1 * te2;
  for (int idx = blockIdx.y * blockDim.x + threadIdx.x; idx < ncell * C; idx += gridDim.y * blockDim.x) {
    const int c = idx % C, cell = idx / C;
    const int x0 = cell % te0, x1 = (cell / te0) % te1, x2 = cell / (te0 * te1);
    const int64_t lc = x0 + (int64_t)T * (x1 + (int64_t)T * x2);
    const float2* src = partial + ((int64_t)it.slot0 * tilecells + lc) * C + c;
    const int64_t stride = tilecells * C;
    float2 acc = make_float2(0.f, 0.f);
    int k = 0;
    for (; k + 4 <= it.nslots; k += 4) {   // four loads in flight, summed in slot order
      const float2 v0 = src[(k + 0) * stride], v1 = src[(k + 1) * stride];
      const float2 v2 = src[(k + 2) * stride], v3 = src[(k + 3) * stride];
      acc.x += v0.x; acc.y += v0.y;
      acc.x += v1.x; acc.y += v1.y;
      acc.x += v2.x; acc.y += v2.y;
      acc.x += v3.x; acc.y += v3.y;
    }
    for (; k < it.nslots; ++k) {
      const float2 v = src[k * stride];
      acc.x += v.x;
      acc.y += v.y;
    }
    const int64_t l = (ts0 + x0) + (int64_t)n0 * ((ts1 + x1) + (int64_t)n1 * (ts2 + x2));
    grid[((int64_t)it.frame * ntot + l) * C + c] = acc;
  }
}

// ---------------------------------------------------------------------------
// The 2D gathers' LDS staging: the tile's extended region, (T + taps - 1)^2
// cells x CG coils at ext[e1 * rowstride + e0 * CG + c], and the item's window
// values (wpsi, 2 taps per sample) and base cells relative to the region
// origin (wrel).  Returns the item's sample count.
template <int CG, int TAPS, int TT, int NTH = 256, bool WIN = true>
__device__ __forceinline__ int stage_region_2d(const SampArgs& a, const int4 it, int T, int E, int rowstride,
                                               int nt0, int cbase, float2* ext, float* wpsi, int* wrel) {
  const int m = TAPS / 2 - 1;
  const int t0 = it.y % nt0, t1 = it.y / nt0;
  const int org0 = t0 * T - m, org1 = t1 * T - m;
  const int64_t gbase = (int64_t)it.x * a.ntot;
  if constexpr (TT > 0) {
    // a thread owns one (column e0, coil) of the region and loads it from all
    // E rows (E loads in flight); the row index is the loop's, so the row
    // address is wave-uniform and the per-load work is one add
    constexpr int EC = (TT + TAPS - 1), NQ = EC * CG;
    for (int q = threadIdx.x; q < NQ; q += NTH) {
      const int e0 = q / CG, c = q - e0 * CG;
      const bool cok = cbase + c < a.C;
      int l0 = org0 + e0;
      if (l0 < 0) l0 += a.n0; else if (l0 >= a.n0) l0 -= a.n0;
      const float2* src = a.grid + (gbase + l0) * a.C + cbase + c;
      float2 v[EC];
#pragma unroll
      for (int e1 = 0; e1 < EC; ++e1) {
        int l1 = org1 + e1;
        if (l1 < 0) l1 += a.n1; else if (l1 >= a.n1) l1 -= a.n1;
        v[e1] = cok ? src[(int64_t)a.n0 * l1 * a.C] : make_float2(0.f, 0.f);
      }
#pragma unroll
      for (int e1 = 0; e1 < EC; ++e1) ext[e1 * rowstride + e0 * CG + c] = v[e1];
    }
  } else {
    // eight global loads in flight per thread before any is waited on
    constexpr int U = 8, STEP = NTH / CG;
    const int c = threadIdx.x % CG;
    const bool cok = cbase + c < a.C;
    for (int e00 = threadIdx.x / CG; e00 < E * E; e00 += STEP * U) {
      float2 v[U];
      int dst[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int e = e00 + u * STEP;
        v[u] = make_float2(0.f, 0.f);
        dst[u] = -1;
        if (e < E * E) {
          const int e1 = e / E, e0 = e - e1 * E;
          int l0 = org0 + e0, l1 = org1 + e1;   // in [-m, n + T + m): one wrap suffices
          if (l0 < 0) l0 += a.n0; else if (l0 >= a.n0) l0 -= a.n0;
          if (l1 < 0) l1 += a.n1; else if (l1 >= a.n1) l1 -= a.n1;
          if (cok) v[u] = a.grid[(gbase + l0 + (int64_t)a.n0 * l1) * a.C + cbase + c];
          dst[u] = e1 * rowstride + e0 * CG + c;
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (dst[u] >= 0) ext[dst[u]] = v[u];
    }
  }
  const int ns = it.w - it.z;
  if constexpr (!WIN) return ns;
  for (int idx = threadIdx.x; idx < ns * 2 * TAPS; idx += blockDim.x)
    wpsi[idx] = a.psi[(int64_t)it.z * 2 * TAPS + idx];
  for (int k = threadIdx.x; k < ns; k += blockDim.x) {
    const int s = it.z + k;
    int r0 = a.base[2 * s] - org0, r1 = a.base[2 * s + 1] - org1;
    if (r0 < 0) r0 += a.n0; else if (r0 >= a.n0) r0 -= a.n0;
    if (r1 < 0) r1 += a.n1; else if (r1 >= a.n1) r1 -= a.n1;
    wrel[2 * k] = r0;
    wrel[2 * k + 1] = r1;
  }
  return ns;
}

// ---------------------------------------------------------------------------
// forward B with the grid staged in LDS (2D).  Work item = (frame, tile,
// samples [sbeg, send) whose base cell lies in the tile); block = 4 waves; one
// coil group of CG coils per blockIdx.y.  The extended region of the tile,
// (T + taps - 1)^2 cells, and the items' window values are staged once; every
// tap is then an LDS read.  Lanes = (coil, tap row slot), one sample per wave
// at a time, a lane sum over the row slots.
template <int CG, int TAPS, int TT>
__global__ void __launch_bounds__(256) k_interp_lds(SampArgs a, const int4* items, int nitems, int T_,
                                                    int rowstride_, int maxs, int nt0) {
  constexpr int RS = 64 / CG;
  extern __shared__ float2 sh[];
  const int item = blockIdx.x;
  const int4 it = items[item];             // frame, tile, sbeg, send
  if (a.fdone && a.fdone[it.x]) return;
  // TT > 0: the tile side as a constant, so that the staging's divisions by the
  // region width are multiplications
  const int T = TT > 0 ? TT : T_;
  const int E = T + TAPS - 1;
  const int rowstride = TT > 0 ? (TT + TAPS - 1) * CG + ((CG == 8) ? 8 : 0) : rowstride_;
  float2* ext = sh;                                          // [E][rowstride]
  float* wpsi = (float*)(sh + (size_t)E * rowstride);        // [maxs][2*TAPS]
  int* wrel = (int*)(wpsi + (size_t)maxs * 2 * TAPS);        // [maxs][2]
  const int cbase = blockIdx.y * CG;
  const int ns = stage_region_2d<CG, TAPS, TT>(a, it, T, E, rowstride, nt0, cbase, ext, wpsi, wrel);
  lds_barrier();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int cl = lane % CG, rs = lane / CG;
  const int c = cbase + cl;
  for (int k = wave; k < ns; k += 4) {
    const float* ps = wpsi + k * 2 * TAPS;
    const int r0 = wrel[2 * k], r1 = wrel[2 * k + 1];
    float2 acc = make_float2(0.f, 0.f);
#pragma unroll
    for (int s1 = rs; s1 < TAPS; s1 += RS) {
      const float2* row = ext + (r1 + s1) * rowstride + r0 * CG + cl;
      float2 racc = make_float2(0.f, 0.f);
#pragma unroll
      for (int s0 = 0; s0 < TAPS; ++s0) {
        const float2 v = row[s0 * CG];
        const float p0 = ps[s0];
        racc.x += p0 * v.x;
        racc.y += p0 * v.y;
      }
      const float wr = ps[TAPS + s1];
      acc.x += wr * racc.x;
      acc.y += wr * racc.y;
    }
    acc = make_float2(class_sum<CG>(acc.x), class_sum<CG>(acc.y));   // over the row slots
    if (rs == 0 && c < a.C) a.y[(int64_t)(it.z + k) * a.C + c] = acc;
  }
}

// The same gather with lanes = (coil, sample): a wave takes 64 / CG samples
// at a time and every lane sums all taps of its sample for its coil, so no
// lane sum is needed (about 0.75 of the VALU instructions per sample of
// k_interp_lds); the lanes of one sample read CG consecutive coils of a region
// row.  Only the region is staged (its LDS, 29 KB at 16 coils, lets four
// 512-thread blocks share a CU); window values and base cells are read from
// global memory, one address per sample (a broadcast to its CG lanes).
// CPL = 2: a lane takes two adjacent coils, so that every tap is one 16-byte
// LDS read (ds_read_b128 moves about three times the bytes per clock of
// ds_read_b64 on gfx950); rowstride and CG are even, so the pairs are aligned.
template <int CG, int TAPS, int TT, int NTH, int CPL>
__global__ void __launch_bounds__(NTH) k_interp_cs(SampArgs a, const int4* items, int nitems, int T_,
                                                   int rowstride_, int maxs, int nt0) {
  static_assert(CPL == 1 || (CPL == 2 && CG % 2 == 0), "one or two coils per lane");
  constexpr int LPS = CG / CPL;            // lanes per sample
  constexpr int SPW = 64 / LPS, NW = NTH / 64;
  extern __shared__ float2 sh[];
  const int item = blockIdx.x;
  const int4 it = items[item];             // frame, tile, sbeg, send
  if (a.fdone && a.fdone[it.x]) return;
  const int T = TT > 0 ? TT : T_;
  const int E = T + TAPS - 1;
  const int rowstride = TT > 0 ? (TT + TAPS - 1) * CG + ((CG == 8) ? 8 : 0) : rowstride_;
  float2* ext = sh;                                          // [E][rowstride]
  const int cbase = blockIdx.y * CG;
  const int ns = stage_region_2d<CG, TAPS, TT, NTH, false>(a, it, T, E, rowstride, nt0, cbase, ext, nullptr,
                                                           nullptr);
  const int m = TAPS / 2 - 1;
  const int org0 = (it.y % nt0) * T - m, org1 = (it.y / nt0) * T - m;
  lds_barrier();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int cl = lane % LPS, sl = lane / LPS;
  const int c = cbase + cl * CPL;
  // the window values and base cell of the lane's next sample are loaded while
  // the current one is summed (lanes past the item's end load sample sbeg)
  float nx[TAPS], ny[TAPS];
  int nb0, nb1;
  auto load_next = [&](int k) {
    const int smp = it.z + (k < ns ? k : 0);
    const float* ps = a.psi + (int64_t)smp * 2 * TAPS;
#pragma unroll
    for (int s0 = 0; s0 < TAPS; ++s0) {
      nx[s0] = ps[s0];
      ny[s0] = ps[TAPS + s0];
    }
    nb0 = a.base[2 * smp];
    nb1 = a.base[2 * smp + 1];
  };
  load_next(wave * SPW + sl);
  for (int k0 = wave * SPW; k0 < ns; k0 += NW * SPW) {
    const int k = k0 + sl;
    float px[TAPS], py[TAPS];
#pragma unroll
    for (int s0 = 0; s0 < TAPS; ++s0) {
      px[s0] = nx[s0];
      py[s0] = ny[s0];
    }
    int r0 = nb0 - org0, r1 = nb1 - org1;
    load_next(k + NW * SPW);
    if (k >= ns) continue;   // the item's last step: lanes past its end stay idle
    const int smp = it.z + k;
    if (r0 < 0) r0 += a.n0; else if (r0 >= a.n0) r0 -= a.n0;
    if (r1 < 0) r1 += a.n1; else if (r1 >= a.n1) r1 -= a.n1;
    if constexpr (CPL == 1) {
      const float2* base = ext + r1 * rowstride + r0 * CG + cl;
      float2 acc = make_float2(0.f, 0.f);
#pragma unroll
      for (int s1 = 0; s1 < TAPS; ++s1) {
        const float2* row = base + s1 * rowstride;
        float2 racc = make_float2(0.f, 0.f);
#pragma unroll
        for (int s0 = 0; s0 < TAPS; ++s0) {
          const float2 v = row[s0 * CG];
          racc.x += px[s0] * v.x;
          racc.y += px[s0] * v.y;
        }
        acc.x += py[s1] * racc.x;
        acc.y += py[s1] * racc.y;
      }
      if (c < a.C) a.y[(int64_t)smp * a.C + c] = acc;
    } else {
      const float4* base = reinterpret_cast<const float4*>(ext + r1 * rowstride + r0 * CG + 2 * cl);
      float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
      for (int s1 = 0; s1 < TAPS; ++s1) {
        const float4* row = base + s1 * (rowstride / 2);
        float4 racc = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int s0 = 0; s0 < TAPS; ++s0) {
          const float4 v = row[s0 * (CG / 2)];
          racc.x += px[s0] * v.x;
          racc.y += px[s0] * v.y;
          racc.z += px[s0] * v.z;
          racc.w += px[s0] * v.w;
        }
        acc.x += py[s1] * racc.x;
        acc.y += py[s1] * racc.y;
        acc.z += py[s1] * racc.z;
        acc.w += py[s1] * racc.w;
      }
      float2* yo = a.y + (int64_t)smp * a.C + c;
      if ((a.C & 1) == 0 && c + 1 < a.C) {   // (smp C + c) even: 16-byte aligned
        *reinterpret_cast<float4*>(yo) = acc;
      } else {
        if (c < a.C) yo[0] = make_float2(acc.x, acc.y);
        if (c + 1 < a.C) yo[1] = make_float2(acc.z, acc.w);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// adjoint B^T, owner computes, visits processed in batches of 64: the batch's
// sample ids, in-tile tap ranges, window values and (weighted) data are first
// loaded cooperatively into LDS, then every wave runs through the batch.
// Tile row x1 belongs to row slot x1 mod (W*RS) (slot = wave*RS + rs), so
// each (cell, coil) has one writer and visits are applied in list
// order: race-free, deterministic, no atomics.  Slots own absolute tile rows
// (row x1 == slot mod W*RS), never footprint-relative rows, so two waves at
// different visits can not reach the same cell.  W*RS must divide T.
template <int CG, int TAPS, int D>
__global__ void __launch_bounds__(256) k_grid_batched(GridArgs a) {
  constexpr int RS = 64 / CG;
  constexpr int B = 64;
  constexpr int NPSI = D * TAPS;
  extern __shared__ float2 sh[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int W = blockDim.x >> 6;
  const int cl = lane % CG, rs = lane / CG;
  const int slot = wave * RS + rs, nslots = W * RS;
  const GridItem it = a.items[blockIdx.x];
  if (a.fdone && a.fdone[it.frame]) return;
  const int T = a.T, rowstride = a.rowstride;
  const int cbase = blockIdx.y * CG;
  const int c = cbase + cl;
  const int nrows_tile = (D == 3) ? T * T : T;
  float2* tile = sh;                                               // [rows][rowstride]
  float2* ybat = tile + (size_t)nrows_tile * rowstride;            // [B][CG]
  float* bpsi = (float*)(ybat + B * CG);                           // [B][NPSI]
  int* bmeta = (int*)(bpsi + B * NPSI);                            // [B][8]: s, rel0..2, lo0,hi0 packed

  int t0 = it.tile % a.nt0, t1 = (it.tile / a.nt0) % a.nt1, t2 = (D == 3) ? it.tile / (a.nt0 * a.nt1) : 0;
  const int ts0 = t0 * T, ts1 = t1 * T, ts2 = t2 * T;
  const int te0 = min(T, a.n0 - ts0), te1 = min(T, a.n1 - ts1), te2 = (D == 3) ? min(T, a.n2 - ts2) : 1;

  for (int idx = threadIdx.x; idx < nrows_tile * rowstride; idx += blockDim.x) tile[idx] = make_float2(0.f, 0.f);

  for (int vb = it.vbeg; vb < it.vend; vb += B) {
    const int nb = min(B, it.vend - vb);
    lds_barrier();
    if (threadIdx.x < nb) {
      const int s = a.visits[vb + threadIdx.x];
      const int* bs = a.base + (int64_t)s * D;
      int rel0 = bs[0] - ts0; if (rel0 < 0) rel0 += a.n0; if (rel0 >= te0) rel0 -= a.n0;
      int rel1 = bs[1] - ts1; if (rel1 < 0) rel1 += a.n1; if (rel1 >= te1) rel1 -= a.n1;
      int rel2 = 0;
      if (D == 3) { rel2 = bs[2] - ts2; if (rel2 < 0) rel2 += a.n2; if (rel2 >= te2) rel2 -= a.n2; }
      int* mt = bmeta + threadIdx.x * 8;
      mt[0] = s;
      mt[1] = rel0; mt[2] = rel1; mt[3] = rel2;
      mt[4] = max(0, -rel0) | (min(TAPS, te0 - rel0) << 16);
      mt[5] = max(0, -rel1) | (min(TAPS, te1 - rel1) << 16);
      mt[6] = (D == 3) ? (max(0, -rel2) | (min(TAPS, te2 - rel2) << 16)) : (0 | (1 << 16));
      const float* ps = a.psi + (int64_t)s * NPSI;
#pragma unroll
      for (int q = 0; q < NPSI; ++q) bpsi[threadIdx.x * NPSI + q] = ps[q];
    }
    lds_barrier();
    for (int idx = threadIdx.x; idx < nb * CG; idx += blockDim.x) {
      const int v = idx / CG, cc = idx % CG;
      const int s = bmeta[v * 8];
      float2 yv = make_float2(0.f, 0.f);
      if (cbase + cc < a.C) {
        yv = a.y[(int64_t)s * a.C + cbase + cc];
        if (a.w) { const float ww = a.w[s]; yv.x *= ww; yv.y *= ww; }
      }
      ybat[idx] = yv;
    }
    lds_barrier();
    for (int v = 0; v < nb; ++v) {
      const int* mt = bmeta + v * 8;
      const int rel0 = mt[1], rel1 = mt[2], rel2 = mt[3];
      const int lo0 = mt[4] & 0xffff, hi0 = mt[4] >> 16;
      const int lo1 = mt[5] & 0xffff, hi1 = mt[5] >> 16;
      const int lo2 = mt[6] & 0xffff, hi2 = mt[6] >> 16;
      const float* ps = bpsi + v * NPSI;
      const float2 yv = ybat[v * CG + cl];
      float p0v[TAPS];
#pragma unroll
      for (int s0 = 0; s0 < TAPS; ++s0) p0v[s0] = ps[s0];
      // this slot owns the tile rows x1 == slot (mod nslots) of every plane
      const int r1a = rel1 + lo1, r1b = rel1 + hi1;
      int first = r1a + ((slot - r1a) % nslots + nslots) % nslots;
      for (int s2 = lo2; s2 < hi2; ++s2) {
        const float w2 = (D == 3) ? ps[2 * TAPS + s2] : 1.f;
        const int plane = (D == 3) ? (rel2 + s2) * T : 0;
        for (int x1 = first; x1 < r1b; x1 += nslots) {
          const int s1 = x1 - rel1;
          const float wr = ps[TAPS + s1] * w2;
          const float vx = yv.x * wr, vy = yv.y * wr;
          float2* dst = tile + (x1 + plane) * rowstride + rel0 * CG + cl;
          // all reads of the row first (distinct cells), then the updates
          float2 cur[TAPS];
#pragma unroll
          for (int s0 = 0; s0 < TAPS; ++s0)
            if (s0 >= lo0 && s0 < hi0) cur[s0] = dst[s0 * CG];
#pragma unroll
          for (int s0 = 0; s0 < TAPS; ++s0)
            if (s0 >= lo0 && s0 < hi0) {
              cur[s0].x += p0v[s0] * vx;
              cur[s0].y += p0v[s0] * vy;
              dst[s0 * CG] = cur[s0];
            }
        }
      }
    }
  }
  lds_barrier();

  const int ncell = te0 * te1 * te2;
  for (int idx = threadIdx.x; idx < ncell * CG; idx += blockDim.x) {
    const int cc = idx % CG, cell = idx / CG;
    const int cg = cbase + cc;
    if (cg >= a.C) continue;
    const int x0 = cell % te0, x1 = (cell / te0) % te1, x2 = cell / (te0 * te1);
    const float2 v = tile[(x1 + x2 * T) * rowstride + x0 * CG + cc];
    if (it.slot < 0) {
      const int64_t l = (ts0 + x0) + (int64_t)a.n0 * ((ts1 + x1) + (int64_t)a.n1 * (ts2 + x2));
      a.grid[((int64_t)it.frame * a.ntot + l) * a.C + cg] = v;
    } else {
      const int64_t lc = x0 + (int64_t)T * (x1 + (int64_t)T * x2);
      a.partial[((int64_t)it.slot * a.tilecells + lc) * a.C + cg] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// adjoint B^T, owner computes, one wave per visit.  Visit records are
// precomputed on the host: {s, rel0 | rel1 << 16, lo0 | hi0 << 8 | lo1 << 16 |
// hi1 << 24, rel2 | lo2 << 16 | hi2 << 24} (rel as int16, in-tile tap ranges
// [lo, hi)).  A batch of 64 records, their window values and (weighted) data
// are staged in LDS by the whole block; the four waves then take disjoint
// visits and add their taps into the LDS tile with LDS float atomics.  The
// summation order inside a tile therefore varies between runs (the result is
// the same to rounding); CGS_DETERMINISTIC plans use k_grid_batched instead.
template <int CG, int TAPS, int D>
__global__ void __launch_bounds__(256) k_grid_atomic(GridArgs a, const int4* __restrict__ vrec) {
  constexpr int RS = 64 / CG;
  constexpr int B = 64;
  constexpr int NPSI = D * TAPS;
  extern __shared__ float2 sh[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int cl = lane % CG, rs = lane / CG;
  const GridItem it = a.items[blockIdx.x];
  if (a.fdone && a.fdone[it.frame]) return;
  const int T = a.T, rowstride = a.rowstride;
  const int cbase = blockIdx.y * CG;
  const int nrows_tile = (D == 3) ? T * T : T;
  float2* tile = sh;                                               // [rows][rowstride]
  float2* ybat = tile + (size_t)nrows_tile * rowstride;            // [B][CG]
  float* bpsi = (float*)(ybat + B * CG);                           // [B][NPSI]
  int4* brec = (int4*)(bpsi + B * NPSI);                           // [B]

  int t0 = it.tile % a.nt0, t1 = (it.tile / a.nt0) % a.nt1, t2 = (D == 3) ? it.tile / (a.nt0 * a.nt1) : 0;
  const int ts0 = t0 * T, ts1 = t1 * T, ts2 = t2 * T;
  const int te0 = min(T, a.n0 - ts0), te1 = min(T, a.n1 - ts1), te2 = (D == 3) ? min(T, a.n2 - ts2) : 1;

  for (int idx = threadIdx.x; idx < nrows_tile * rowstride; idx += blockDim.x) tile[idx] = make_float2(0.f, 0.f);

  for (int vb = it.vbeg; vb < it.vend; vb += B) {
    const int nb = min(B, it.vend - vb);
    lds_barrier();
    if (threadIdx.x < nb) brec[threadIdx.x] = vrec[vb + threadIdx.x];
    lds_barrier();
    {
      // window values: NPSI * nb <= 64 * 24 floats, 6 per thread at most
      float pv[(B * NPSI + 255) / 256];
#pragma unroll
      for (int u = 0; u < (B * NPSI + 255) / 256; ++u) {
        const int idx = threadIdx.x + u * 256;
        if (idx < nb * NPSI) {
          const int v = idx / NPSI, q = idx - v * NPSI;
          pv[u] = a.psi[(int64_t)brec[v].x * NPSI + q];
        }
      }
      // data: nb * CG <= 2048 float2, 8 per thread at most
      float2 yv[(B * CG + 255) / 256];
#pragma unroll
      for (int u = 0; u < (B * CG + 255) / 256; ++u) {
        const int idx = threadIdx.x + u * 256;
        yv[u] = make_float2(0.f, 0.f);
        if (idx < nb * CG) {
          const int v = idx / CG, cc = idx - v * CG;
          const int s = brec[v].x;
          if (cbase + cc < a.C) {
            yv[u] = a.y[(int64_t)s * a.C + cbase + cc];
            if (a.w) { const float ww = a.w[s]; yv[u].x *= ww; yv[u].y *= ww; }
          }
        }
      }
#pragma unroll
      for (int u = 0; u < (B * NPSI + 255) / 256; ++u) {
        const int idx = threadIdx.x + u * 256;
        if (idx < nb * NPSI) bpsi[idx] = pv[u];
      }
#pragma unroll
      for (int u = 0; u < (B * CG + 255) / 256; ++u) {
        const int idx = threadIdx.x + u * 256;
        if (idx < nb * CG) ybat[idx] = yv[u];
      }
    }
    lds_barrier();
    for (int v = wave; v < nb; v += 4) {
      const int4 r = brec[v];
      const int rel0 = (int)(short)(r.y & 0xffff), rel1 = (int)(short)(r.y >> 16);
      const int lo0 = r.z & 0xff, hi0 = (r.z >> 8) & 0xff, lo1 = (r.z >> 16) & 0xff, hi1 = (r.z >> 24) & 0xff;
      const int rel2 = (D == 3) ? (int)(short)(r.w & 0xffff) : 0;
      const int lo2 = (D == 3) ? ((r.w >> 16) & 0xff) : 0, hi2 = (D == 3) ? ((r.w >> 24) & 0xff) : 1;
      const float* ps = bpsi + v * NPSI;
      float p0v[TAPS];
#pragma unroll
      for (int s0 = 0; s0 < TAPS; ++s0) p0v[s0] = ps[s0];
      const float2 yv = ybat[v * CG + cl];
      for (int s2 = lo2; s2 < hi2; ++s2) {
        const float w2 = (D == 3) ? ps[2 * TAPS + s2] : 1.f;
        const int plane = (D == 3) ? (rel2 + s2) * T : 0;
        for (int s1 = lo1 + rs; s1 < hi1; s1 += RS) {
          const float wr = ps[TAPS + s1] * w2;
          const float vx = yv.x * wr, vy = yv.y * wr;
          float* dst = (float*)(tile + (rel1 + s1 + plane) * rowstride + rel0 * CG + cl);
#pragma unroll
          for (int s0 = 0; s0 < TAPS; ++s0)
            if (s0 >= lo0 && s0 < hi0) {
              atomicAdd(dst + 2 * s0 * CG, p0v[s0] * vx);
              atomicAdd(dst + 2 * s0 * CG + 1, p0v[s0] * vy);
            }
        }
      }
    }
  }
  lds_barrier();

  const int ncell = te0 * te1 * te2;
  for (int idx = threadIdx.x; idx < ncell * CG; idx += blockDim.x) {
    const int cc = idx % CG, cell = idx / CG;
    const int cg = cbase + cc;
    if (cg >= a.C) continue;
    const int x0 = cell % te0, x1 = (cell / te0) % te1, x2 = cell / (te0 * te1);
    const float2 v = tile[(x1 + x2 * T) * rowstride + x0 * CG + cc];
    if (it.slot < 0) {
      const int64_t l = (ts0 + x0) + (int64_t)a.n0 * ((ts1 + x1) + (int64_t)a.n1 * (ts2 + x2));
      a.grid[((int64_t)it.frame * a.ntot + l) * a.C + cg] = v;
    } else {
      const int64_t lc = x0 + (int64_t)T * (x1 + (int64_t)T * x2);
      a.partial[((int64_t)it.slot * a.tilecells + lc) * a.C + cg] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// adjoint B^T in registers (2D, 8x8 tiles): one wave owns one work item (a
// tile or a chunk of a heavy tile's visit list); lane = grid cell of the tile
// and its CGR coil accumulators live in VGPRs.  Batches of 32 visit records,
// window values and weighted data are staged in a wave-private LDS region and
// read back as broadcasts, so a visit costs one weight per lane and 2*CGR FMAs.
// Each cell has one owner and visits are applied in list order: no atomics,
// no LDS read-modify-write, deterministic.
template <int CGR, int TAPS>
__global__ void __launch_bounds__(256) k_grid_reg(GridArgs a, const int4* __restrict__ vrec, int nitems) {
  constexpr int B = 32, NPSI = 2 * TAPS;
  constexpr int WAVE_FLOATS = B * 2 * CGR + B * NPSI + B * 4;
  extern __shared__ float shf[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int item = blockIdx.x * 4 + wave;
  if (item >= nitems) return;          // waves of a block never synchronise with each other
  const GridItem it = a.items[item];
  if (a.fdone && a.fdone[it.frame]) return;
  float* wy = shf + wave * WAVE_FLOATS;           // [B][2*CGR]
  float* wpsi = wy + B * 2 * CGR;                 // [B][NPSI]
  int4* wrec = (int4*)(wpsi + B * NPSI);          // [B]
  const int cbase = blockIdx.y * CGR;
  const int t0 = it.tile % a.nt0, t1 = it.tile / a.nt0;
  const int ts0 = t0 * 8, ts1 = t1 * 8;
  const int te0 = min(8, a.n0 - ts0), te1 = min(8, a.n1 - ts1);
  const int x0 = lane & 7, x1 = lane >> 3;
  float acc[2 * CGR];
#pragma unroll
  for (int k = 0; k < 2 * CGR; ++k) acc[k] = 0.f;

  for (int vb = it.vbeg; vb < it.vend; vb += B) {
    const int nb = min(B, it.vend - vb);
    if (lane < nb) wrec[lane] = vrec[vb + lane];
    __builtin_amdgcn_wave_barrier();
    {
      float pv[(B * NPSI + 63) / 64];
#pragma unroll
      for (int u = 0; u < (B * NPSI + 63) / 64; ++u) {
        const int idx = lane + 64 * u;
        if (idx < nb * NPSI) {
          const int v = idx / NPSI, q = idx - v * NPSI;
          pv[u] = a.psi[(int64_t)wrec[v].x * NPSI + q];
        }
      }
      float2 yv[(B * CGR + 63) / 64];
#pragma unroll
      for (int u = 0; u < (B * CGR + 63) / 64; ++u) {
        const int idx = lane + 64 * u;
        yv[u] = make_float2(0.f, 0.f);
        if (idx < nb * CGR) {
          const int v = idx / CGR, c = idx - v * CGR;
          const int s = wrec[v].x;
          if (cbase + c < a.C) {
            yv[u] = a.y[(int64_t)s * a.C + cbase + c];
            if (a.w) { const float ww = a.w[s]; yv[u].x *= ww; yv[u].y *= ww; }
          }
        }
      }
#pragma unroll
      for (int u = 0; u < (B * NPSI + 63) / 64; ++u) {
        const int idx = lane + 64 * u;
        if (idx < nb * NPSI) wpsi[idx] = pv[u];
      }
#pragma unroll
      for (int u = 0; u < (B * CGR + 63) / 64; ++u) {
        const int idx = lane + 64 * u;
        if (idx < nb * CGR) ((float2*)wy)[idx] = yv[u];
      }
    }
    __builtin_amdgcn_wave_barrier();
    for (int v = 0; v < nb; ++v) {
      const int4 r = wrec[v];
      const int dx = x0 - (int)(short)(r.y & 0xffff), dy = x1 - (int)(short)(r.y >> 16);
      float wgt = 0.f;
      if (dx >= 0 && dx < TAPS && dy >= 0 && dy < TAPS) wgt = wpsi[v * NPSI + dx] * wpsi[v * NPSI + TAPS + dy];
      const float4* yy = (const float4*)(wy + v * 2 * CGR);
#pragma unroll
      for (int k = 0; k < CGR / 2; ++k) {
        const float4 t = yy[k];
        acc[4 * k + 0] += wgt * t.x;
        acc[4 * k + 1] += wgt * t.y;
        acc[4 * k + 2] += wgt * t.z;
        acc[4 * k + 3] += wgt * t.w;
      }
    }
    __builtin_amdgcn_wave_barrier();
  }

  if (x0 < te0 && x1 < te1) {
    float2* out;
    if (it.slot < 0) {
      const int64_t l = (ts0 + x0) + (int64_t)a.n0 * (ts1 + x1);
      out = a.grid + ((int64_t)it.frame * a.ntot + l) * a.C + cbase;
    } else {
      out = a.partial + ((int64_t)it.slot * a.tilecells + x0 + 8 * x1) * a.C + cbase;
    }
#pragma unroll
    for (int c = 0; c < CGR; ++c)
      if (cbase + c < a.C) out[c] = make_float2(acc[2 * c], acc[2 * c + 1]);
  }
}

// 3D: one block of 8 waves per work item, wave w owning plane w of the
// 8 x 8 x 8 tile (lane = cell of the plane, CGR coil accumulators in VGPRs).
// Visit batches (records, window values, weighted data) are staged once per
// block in LDS; a wave skips the visits whose footprint misses its plane.
template <int CGR, int TAPS>
__global__ void __launch_bounds__(512) k_grid_reg3(GridArgs a, const int4* __restrict__ vrec, int nitems) {
  constexpr int B = 64, NPSI = 3 * TAPS;
  __shared__ float2 sy[B * CGR];
  __shared__ float spsi[B * NPSI];
  __shared__ int4 srec[B];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int item = blockIdx.x;
  const GridItem it = a.items[item];
  if (a.fdone && a.fdone[it.frame]) return;   // block-uniform
  const int cbase = blockIdx.y * CGR;
  const int t0 = it.tile % a.nt0, t1 = (it.tile / a.nt0) % a.nt1, t2 = it.tile / (a.nt0 * a.nt1);
  const int ts0 = t0 * 8, ts1 = t1 * 8, ts2 = t2 * 8;
  const int te0 = min(8, a.n0 - ts0), te1 = min(8, a.n1 - ts1), te2 = min(8, a.n2 - ts2);
  const int x0 = lane & 7, x1 = lane >> 3;
  float acc[2 * CGR];
#pragma unroll
  for (int k = 0; k < 2 * CGR; ++k) acc[k] = 0.f;

  for (int vb = it.vbeg; vb < it.vend; vb += B) {
    const int nb = min(B, it.vend - vb);
    lds_barrier();   // the previous batch is consumed
    if (tid < nb) srec[tid] = vrec[vb + tid];
    lds_barrier();
    for (int idx = tid; idx < nb * NPSI; idx += 512) {
      const int v = idx / NPSI, q = idx - v * NPSI;
      spsi[idx] = a.psi[(int64_t)srec[v].x * NPSI + q];
    }
    for (int idx = tid; idx < nb * CGR; idx += 512) {
      const int v = idx / CGR, c = idx - v * CGR;
      const int smp = srec[v].x;
      float2 yv = make_float2(0.f, 0.f);
      if (cbase + c < a.C) {
        yv = a.y[(int64_t)smp * a.C + cbase + c];
        if (a.w) { const float ww = a.w[smp]; yv.x *= ww; yv.y *= ww; }
      }
      sy[idx] = yv;
    }
    lds_barrier();
    for (int v = 0; v < nb; ++v) {
      const int4 r = srec[v];
      const int dz = w - (int)(short)(r.w & 0xffff);
      if (dz < 0 || dz >= TAPS) continue;          // uniform per wave
      const int dx = x0 - (int)(short)(r.y & 0xffff), dy = x1 - (int)(short)(r.y >> 16);
      float wgt = 0.f;
      if (dx >= 0 && dx < TAPS && dy >= 0 && dy < TAPS)
        wgt = spsi[v * NPSI + dx] * spsi[v * NPSI + TAPS + dy] * spsi[v * NPSI + 2 * TAPS + dz];
      const float4* yy = (const float4*)(sy + v * CGR);
#pragma unroll
      for (int k = 0; k < CGR / 2; ++k) {
        const float4 t = yy[k];
        acc[4 * k + 0] += wgt * t.x;
        acc[4 * k + 1] += wgt * t.y;
        acc[4 * k + 2] += wgt * t.z;
        acc[4 * k + 3] += wgt * t.w;
      }
    }
  }

  if (x0 < te0 && x1 < te1 && w < te2) {
    float2* out;
    if (it.slot < 0) {
      const int64_t l = (ts0 + x0) + (int64_t)a.n0 * ((ts1 + x1) + (int64_t)a.n1 * (ts2 + w));
      out = a.grid + ((int64_t)it.frame * a.ntot + l) * a.C + cbase;
    } else {
      out = a.partial + ((int64_t)it.slot * a.tilecells + x0 + 8 * (x1 + 8 * w)) * a.C + cbase;
    }
#pragma unroll
    for (int c = 0; c < CGR; ++c)
      if (cbase + c < a.C) out[c] = make_float2(acc[2 * c], acc[2 * c + 1]);
  }
}

// 3D on the matrix cores: one block of 8 waves per work item, wave w owning
// plane w of the 8 x 8 x 8 tile as D[64 cells][32 columns] (16 coils, re/im).
// Visit batches (records, window values, weighted data) are staged once per
// block; each wave compacts the visits whose footprint meets its plane (ballot
// and prefix count) and applies them two at a time with v_mfma_f32_32x32x2f32
// per 32-cell half.  Visits keep their list order within a wave:
// deterministic, no atomics.
template <int TAPS>
__global__ void __launch_bounds__(512) k_grid_mfma3(GridArgs a, const int4* __restrict__ vrec, int nitems) {
  constexpr int B = 64, NPSI = 3 * TAPS, CC = 16;
  // one more visit slot, B: a footprint outside every tile with zero data,
  // for the odd lane of a wave's last pair (so the pair loop has no branch)
  __shared__ float2 sy[(B + 1) * CC];
  __shared__ float spsi[(B + 1) * NPSI];
  __shared__ int4 srec[B + 1];
  __shared__ unsigned char slist[8][B];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const GridItem it = a.items[blockIdx.x];
  if (a.fdone && a.fdone[it.frame]) return;   // block-uniform
  const int cbase = blockIdx.y * CC;
  const int t0 = it.tile % a.nt0, t1 = (it.tile / a.nt0) % a.nt1, t2 = it.tile / (a.nt0 * a.nt1);
  const int ts0 = t0 * 8, ts1 = t1 * 8, ts2 = t2 * 8;
  const int te0 = min(8, a.n0 - ts0), te1 = min(8, a.n1 - ts1), te2 = min(8, a.n2 - ts2);
  const int i = lane & 31, kk = lane >> 5;
  const int cx = i & 7, cy0 = i >> 3, cy1 = 4 + (i >> 3);
  gf32x16 acc0, acc1;
#pragma unroll
  for (int r = 0; r < 16; ++r) { acc0[r] = 0.f; acc1[r] = 0.f; }
  if (tid < CC) sy[B * CC + tid] = make_float2(0.f, 0.f);
  if (tid < NPSI) spsi[B * NPSI + tid] = 0.f;
  if (tid == 0) srec[B] = make_int4(0, 0x40004000, 0x4000, 0x4000);

  for (int vb = it.vbeg; vb < it.vend; vb += B) {
    const int nb = min(B, it.vend - vb);
    lds_barrier();   // the previous batch is consumed (and the pad slot written)
    if (tid < nb) srec[tid] = vrec[vb + tid];
    lds_barrier();
    {
      // every load of the batch in flight before the first LDS write
      constexpr int UP = (B * NPSI + 511) / 512, UY = (B * CC + 511) / 512;
      float pv[UP];
      float2 yv[UY];
#pragma unroll
      for (int u = 0; u < UP; ++u) {
        const int idx = tid + 512 * u;
        pv[u] = 0.f;
        if (idx < nb * NPSI) {
          const int v = idx / NPSI, q = idx - v * NPSI;
          pv[u] = a.psi[(int64_t)srec[v].x * NPSI + q];
        }
      }
#pragma unroll
      for (int u = 0; u < UY; ++u) {
        const int idx = tid + 512 * u;
        yv[u] = make_float2(0.f, 0.f);
        if (idx < nb * CC) {
          const int v = idx / CC, c = idx - v * CC;
          if (cbase + c < a.C) {
            const int smp = srec[v].x;
            yv[u] = a.y[(int64_t)smp * a.C + cbase + c];
            if (a.w) { const float ww = a.w[smp]; yv[u].x *= ww; yv[u].y *= ww; }
          }
        }
      }
#pragma unroll
      for (int u = 0; u < UP; ++u) {
        const int idx = tid + 512 * u;
        if (idx < nb * NPSI) spsi[idx] = pv[u];
      }
#pragma unroll
      for (int u = 0; u < UY; ++u) {
        const int idx = tid + 512 * u;
        if (idx < nb * CC) sy[idx] = yv[u];
      }
    }
    // the visits of this batch whose footprint meets plane w, in list order
    bool hit = false;
    if (lane < nb) {
      const int dz = w - (int)(short)(srec[lane].w & 0xffff);
      hit = (unsigned)dz < (unsigned)TAPS;
    }
    const unsigned long long mask = __ballot(hit);
    const int cnt = __popcll(mask);
    if (hit) slist[w][__popcll(mask & ((1ull << lane) - 1ull))] = (unsigned char)lane;
    lds_barrier();   // staged data and this wave's list are visible
    for (int p = 0; p < cnt; p += 2) {
      // branch-free: past cnt the pad slot B (zero data, weights selected to 0)
      const int v = (p + kk < cnt) ? slist[w][p + kk] : B;
      const int4 rc = srec[v];
      const int rel0 = (int)(short)(rc.y & 0xffff), rel1 = (int)(short)(rc.y >> 16);
      const int dz = w - (int)(short)(rc.w & 0xffff);   // in [0, TAPS) for listed visits
      const float* ps = spsi + v * NPSI;
      const int dx = cx - rel0, dy0 = cy0 - rel1, dy1 = cy1 - rel1;
      const bool okx = (unsigned)dx < (unsigned)TAPS && (unsigned)dz < (unsigned)TAPS;
      const bool ok0 = okx && (unsigned)dy0 < (unsigned)TAPS, ok1 = okx && (unsigned)dy1 < (unsigned)TAPS;
      const float px = ps[okx ? dx : 0], pz = ps[2 * TAPS + (okx ? dz : 0)];
      const float q0 = ps[TAPS + (ok0 ? dy0 : 0)], q1 = ps[TAPS + (ok1 ? dy1 : 0)];
      const float bv = ((const float*)sy)[v * 2 * CC + i];
      const float pxz = px * pz;
      const float w0 = ok0 ? pxz * q0 : 0.f, w1 = ok1 ? pxz * q1 : 0.f;
      acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(w0, bv, acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(w1, bv, acc1, 0, 0, 0);
    }
  }

  if (w >= te2) return;
  const int ncol = 2 * min(CC, a.C - cbase);
  if (i >= ncol) return;
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int cell = 32 * h + 8 * (r >> 2) + 4 * kk + (r & 3);
      const int x0 = cell & 7, x1 = cell >> 3;
      if (x0 >= te0 || x1 >= te1) continue;
      float* out;
      if (it.slot < 0) {
        const int64_t l = (ts0 + x0) + (int64_t)a.n0 * ((ts1 + x1) + (int64_t)a.n1 * (ts2 + w));
        out = (float*)(a.grid + ((int64_t)it.frame * a.ntot + l) * a.C + cbase);
      } else {
        out = (float*)(a.partial + ((int64_t)it.slot * a.tilecells + x0 + 8 * (x1 + 8 * w)) * a.C + cbase);
      }
      out[i] = h == 0 ? acc0[r] : acc1[r];
    }
}

void launch_grid_mfma3(const GridArgs& a, const int4* vrec, int nitems, hipStream_t s) {
  if (nitems == 0) return;
  const dim3 g((unsigned)nitems, (unsigned)((a.C + 15) / 16));
  switch (a.taps) {
    case 4: hipLaunchKernelGGL((k_grid_mfma3<4>), g, dim3(512), 0, s, a, vrec, nitems); break;
    case 6: hipLaunchKernelGGL((k_grid_mfma3<6>), g, dim3(512), 0, s, a, vrec, nitems); break;
    case 8: hipLaunchKernelGGL((k_grid_mfma3<8>), g, dim3(512), 0, s, a, vrec, nitems); break;
    case 10: hipLaunchKernelGGL((k_grid_mfma3<10>), g, dim3(512), 0, s, a, vrec, nitems); break;
    default: break;
  }
}

void launch_grid_reg3(const GridArgs& a, const int4* vrec, int cgr, int nitems, int ncgroups, hipStream_t s) {
  if (nitems == 0) return;
  const dim3 g((unsigned)nitems, (unsigned)ncgroups);
#define CGS_GR3(CG, TP) hipLaunchKernelGGL((k_grid_reg3<CG, TP>), g, dim3(512), 0, s, a, vrec, nitems)
  if (cgr == 8) {
    switch (a.taps) { case 4: CGS_GR3(8, 4); break; case 6: CGS_GR3(8, 6); break;
                      case 8: CGS_GR3(8, 8); break; case 10: CGS_GR3(8, 10); break; default: break; }
  } else {
    switch (a.taps) { case 4: CGS_GR3(16, 4); break; case 6: CGS_GR3(16, 6); break;
                      case 8: CGS_GR3(16, 8); break; case 10: CGS_GR3(16, 10); break; default: break; }
  }
#undef CGS_GR3
}

template <int CGR>
static void grid_reg_taps(const GridArgs& a, const int4* vrec, int nitems, dim3 g, size_t shm, hipStream_t s) {
  switch (a.taps) {
    case 4: hipLaunchKernelGGL((k_grid_reg<CGR, 4>), g, dim3(256), shm, s, a, vrec, nitems); break;
    case 6: hipLaunchKernelGGL((k_grid_reg<CGR, 6>), g, dim3(256), shm, s, a, vrec, nitems); break;
    case 8: hipLaunchKernelGGL((k_grid_reg<CGR, 8>), g, dim3(256), shm, s, a, vrec, nitems); break;
    case 10: hipLaunchKernelGGL((k_grid_reg<CGR, 10>), g, dim3(256), shm, s, a, vrec, nitems); break;
    default: break;
  }
}

void launch_grid_reg(const GridArgs& a, const int4* vrec, int cgr, int nitems, int ncgroups, hipStream_t s) {
  if (nitems == 0) return;
  const dim3 g((unsigned)((nitems + 3) / 4), (unsigned)ncgroups);
  const size_t shm = sizeof(float) * 4 * (32 * 2 * cgr + 32 * 2 * a.taps + 32 * 4);
  if (cgr == 8) grid_reg_taps<8>(a, vrec, nitems, g, shm, s);
  else if (cgr == 16) grid_reg_taps<16>(a, vrec, nitems, g, shm, s);
  else grid_reg_taps<32>(a, vrec, nitems, g, shm, s);
}

// ---------------------------------------------------------------------------
// adjoint B^T on the matrix cores (2D, 8x8 tiles).  The tile's grid values are
// D[cell][col] (64 cells x the real columns re/im of CC = 16 NT coils) and the
// work item's visit list is a K dimension: D += W[cell][visit] Y[visit][col],
// with W the visit's separable window weight at the cell (0 outside its
// footprint) and Y its weighted data.  One wave per work item; every
// v_mfma_f32_32x32x2f32 (FP32 products and sums) takes two visits for one
// (32-cell half, 32-column tile).  The zero weights of cells outside a visit's
// footprint cost matrix-core cycles instead of VALU issue, and the weights are
// formed while the previous products run.  Operand layout: A lane l =
// W[l % 32][l / 32], B lane l = Y[l / 32][l % 32], accumulator register r of
// lane l = D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].  Each cell still has one
// owner and visits are applied in list order: deterministic, no atomics.

template <int NT, int TAPS, int B>
__global__ void __launch_bounds__(256) k_grid_mfma(GridArgs a, const int4* __restrict__ vrec, int nitems) {
  constexpr int NPSI = 2 * TAPS, NCOL = 32 * NT, CC = 16 * NT;
  constexpr int WAVE_FLOATS = B * NCOL + B * NPSI + B * 4;
  extern __shared__ float shf[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int item = blockIdx.x * 4 + wave;
  if (item >= nitems) return;          // waves of a block never synchronise with each other
  const GridItem it = a.items[item];
  if (a.fdone && a.fdone[it.frame]) return;
  float* wy = shf + wave * WAVE_FLOATS;           // [B][NCOL]
  float* wpsi = wy + B * NCOL;                    // [B][NPSI]
  int4* wrec = (int4*)(wpsi + B * NPSI);          // [B]
  const int cbase = blockIdx.y * CC;
  const int t0 = it.tile % a.nt0, t1 = it.tile / a.nt0;
  const int ts0 = t0 * 8, ts1 = t1 * 8;
  const int te0 = min(8, a.n0 - ts0), te1 = min(8, a.n1 - ts1);
  const int i = lane & 31, kk = lane >> 5;
  const int cx = i & 7, cy0 = i >> 3, cy1 = 4 + (i >> 3);   // A rows: cells i and 32 + i
  gf32x16 acc[2][NT];
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[h][t][r] = 0.f;

  for (int vb = it.vbeg; vb < it.vend; vb += B) {
    const int nb = min(B, it.vend - vb);
    // visits past nb (the odd one of the last pair) get a footprint outside
    // every tile and zero data, so the pair loop below needs no branch
    if (lane < B) wrec[lane] = (lane < nb) ? vrec[vb + lane] : make_int4(0, 0x40004000, 0, 0);
    __builtin_amdgcn_wave_barrier();
    {
      // every load of the batch in flight before the first LDS write
      constexpr int UP = (B * NPSI + 63) / 64, UY = (B * CC + 63) / 64;
      float pv[UP];
      float2 yv[UY];
#pragma unroll
      for (int u = 0; u < UP; ++u) {
        const int idx = lane + 64 * u;
        pv[u] = 0.f;
        if (idx < nb * NPSI) {
          const int v = idx / NPSI, q = idx - v * NPSI;
          pv[u] = a.psi[(int64_t)wrec[v].x * NPSI + q];
        }
      }
#pragma unroll
      for (int u = 0; u < UY; ++u) {   // float2 units: (visit, coil)
        const int idx = lane + 64 * u;
        yv[u] = make_float2(0.f, 0.f);
        if (idx < nb * CC) {
          const int v = idx / CC, c = idx - v * CC;
          if (cbase + c < a.C) {
            const int smp = wrec[v].x;
            yv[u] = a.y[(int64_t)smp * a.C + cbase + c];
            if (a.w) { const float ww = a.w[smp]; yv[u].x *= ww; yv[u].y *= ww; }
          }
        }
      }
#pragma unroll
      for (int u = 0; u < UP; ++u) {
        const int idx = lane + 64 * u;
        if (idx < nb * NPSI) wpsi[idx] = pv[u];
      }
#pragma unroll
      for (int u = 0; u < UY; ++u) {
        const int idx = lane + 64 * u;
        if (idx < B * CC) ((float2*)wy)[idx] = yv[u];   // zeros past nb
      }
    }
    __builtin_amdgcn_wave_barrier();
    // branch-free: the four LDS reads of a pair depend only on its record, and
    // out-of-footprint weights are selected to zero (reads stay in the batch)
    for (int v = 0; v < nb; v += 2) {
      const int vk = v + kk;   // this lane's visit of the pair (K index); vk <= nb < B or vk < B
      const int4 rc = wrec[vk];
      const int rel0 = (int)(short)(rc.y & 0xffff), rel1 = (int)(short)(rc.y >> 16);
      const float* ps = wpsi + vk * NPSI;
      const int dx = cx - rel0, dy0 = cy0 - rel1, dy1 = cy1 - rel1;
      const bool okx = (unsigned)dx < (unsigned)TAPS;
      const bool ok0 = okx && (unsigned)dy0 < (unsigned)TAPS, ok1 = okx && (unsigned)dy1 < (unsigned)TAPS;
      const float px = ps[okx ? dx : 0];
      const float q0 = ps[TAPS + (ok0 ? dy0 : 0)], q1 = ps[TAPS + (ok1 ? dy1 : 0)];
      float bv[NT];
#pragma unroll
      for (int t = 0; t < NT; ++t) bv[t] = wy[vk * NCOL + 32 * t + i];
      const float w0 = ok0 ? px * q0 : 0.f, w1 = ok1 ? px * q1 : 0.f;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        acc[0][t] = __builtin_amdgcn_mfma_f32_32x32x2f32(w0, bv[t], acc[0][t], 0, 0, 0);
        acc[1][t] = __builtin_amdgcn_mfma_f32_32x32x2f32(w1, bv[t], acc[1][t], 0, 0, 0);
      }
    }
    __builtin_amdgcn_wave_barrier();
  }

  const int ncol = 2 * min(CC, a.C - cbase);   // real columns of this coil group
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int cell = 32 * h + 8 * (r >> 2) + 4 * kk + (r & 3);
      const int x0 = cell & 7, x1 = cell >> 3;
      if (x0 >= te0 || x1 >= te1) continue;
      float* out;
      if (it.slot < 0) {
        const int64_t l = (ts0 + x0) + (int64_t)a.n0 * (ts1 + x1);
        out = (float*)(a.grid + ((int64_t)it.frame * a.ntot + l) * a.C + cbase);
      } else {
        out = (float*)(a.partial + ((int64_t)it.slot * a.tilecells + x0 + 8 * x1) * a.C + cbase);
      }
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const int col = 32 * t + i;
        if (col < ncol) out[col] = acc[h][t][r];
      }
    }
}

template <int NT, int B>
static void grid_mfma_taps(const GridArgs& a, const int4* vrec, int nitems, dim3 g, size_t shm, hipStream_t s) {
  switch (a.taps) {
    case 4: hipLaunchKernelGGL((k_grid_mfma<NT, 4, B>), g, dim3(256), shm, s, a, vrec, nitems); break;
    case 6: hipLaunchKernelGGL((k_grid_mfma<NT, 6, B>), g, dim3(256), shm, s, a, vrec, nitems); break;
    case 8: hipLaunchKernelGGL((k_grid_mfma<NT, 8, B>), g, dim3(256), shm, s, a, vrec, nitems); break;
    case 10: hipLaunchKernelGGL((k_grid_mfma<NT, 10, B>), g, dim3(256), shm, s, a, vrec, nitems); break;
    default: break;
  }
}

// NT column tiles of 16 coils per wave (CGS_GRID_MFMA_NT, default 1: more
// waves per SIMD), visit batches of B (CGS_GRID_MFMA_B, 16 or 32)
void launch_grid_mfma(const GridArgs& a, const int4* vrec, int nitems, hipStream_t s) {
  if (nitems == 0) return;
  static const int nt_env = [] { const char* e = std::getenv("CGS_GRID_MFMA_NT"); return e ? std::atoi(e) : 1; }();
  static const int b_env = [] { const char* e = std::getenv("CGS_GRID_MFMA_B"); return e ? std::atoi(e) : 32; }();
  const int NT = (a.C <= 16 || nt_env != 2) ? 1 : 2;
  const int B = b_env == 16 ? 16 : 32;
  const int CC = 16 * NT;
  const dim3 g((unsigned)((nitems + 3) / 4), (unsigned)((a.C + CC - 1) / CC));
  const int npsi = 2 * a.taps;
  const size_t shm = sizeof(float) * 4 * (B * 32 * NT + B * npsi + B * 4);
  if (NT == 1 && B == 16) grid_mfma_taps<1, 16>(a, vrec, nitems, g, shm, s);
  else if (NT == 1) grid_mfma_taps<1, 32>(a, vrec, nitems, g, shm, s);
  else if (B == 16) grid_mfma_taps<2, 16>(a, vrec, nitems, g, shm, s);
  else grid_mfma_taps<2, 32>(a, vrec, nitems, g, shm, s);
}

size_t grid_atomic_lds(int d, int T, int taps, int CG, int rowstride) {
  const int nrows_tile = (d == 3) ? T * T : T;
  return sizeof(float2) * ((size_t)nrows_tile * rowstride + 64 * CG) + sizeof(float) * 64 * d * taps +
         sizeof(int4) * 64;
}

template <int CG, int D>
static void grid_atomic_taps(const GridArgs& a, const int4* vrec, dim3 g, size_t shm, hipStream_t s) {
  switch (a.taps) {
    case 4: hipLaunchKernelGGL((k_grid_atomic<CG, 4, D>), g, dim3(256), shm, s, a, vrec); break;
    case 6: hipLaunchKernelGGL((k_grid_atomic<CG, 6, D>), g, dim3(256), shm, s, a, vrec); break;
    case 8: hipLaunchKernelGGL((k_grid_atomic<CG, 8, D>), g, dim3(256), shm, s, a, vrec); break;
    case 10: hipLaunchKernelGGL((k_grid_atomic<CG, 10, D>), g, dim3(256), shm, s, a, vrec); break;
    default: break;
  }
}

void launch_grid_atomic(const GridArgs& a, const int4* vrec, int cg, int nitems, int ncgroups, hipStream_t s) {
  if (nitems == 0) return;
  const dim3 g((unsigned)nitems, (unsigned)ncgroups);
  const size_t shm = grid_atomic_lds(a.d, a.T, a.taps, cg, a.rowstride);
  if (a.d == 2) {
    if (cg == 8) grid_atomic_taps<8, 2>(a, vrec, g, shm, s);
    else if (cg == 16) grid_atomic_taps<16, 2>(a, vrec, g, shm, s);
    else grid_atomic_taps<32, 2>(a, vrec, g, shm, s);
  } else {
    if (cg == 8) grid_atomic_taps<8, 3>(a, vrec, g, shm, s);
    else if (cg == 16) grid_atomic_taps<16, 3>(a, vrec, g, shm, s);
    else grid_atomic_taps<32, 3>(a, vrec, g, shm, s);
  }
}

size_t grid_batched_lds(int d, int T, int taps, int CG, int rowstride) {
  const int nrows_tile = (d == 3) ? T * T : T;
  return sizeof(float2) * ((size_t)nrows_tile * rowstride + 64 * CG) + sizeof(float) * 64 * d * taps +
         sizeof(int) * 64 * 8;
}

template <int CG, int D>
static void grid_batched_taps(const GridArgs& a, dim3 g, dim3 b, size_t shm, hipStream_t s) {
  switch (a.taps) {
    case 4: hipLaunchKernelGGL((k_grid_batched<CG, 4, D>), g, b, shm, s, a); break;
    case 6: hipLaunchKernelGGL((k_grid_batched<CG, 6, D>), g, b, shm, s, a); break;
    case 8: hipLaunchKernelGGL((k_grid_batched<CG, 8, D>), g, b, shm, s, a); break;
    case 10: hipLaunchKernelGGL((k_grid_batched<CG, 10, D>), g, b, shm, s, a); break;
    default: break;
  }
}

void launch_grid_batched(const GridArgs& a, int cg, int waves, int nitems, int ncgroups, hipStream_t s) {
  if (nitems == 0) return;
  const dim3 g((unsigned)nitems, (unsigned)ncgroups), b((unsigned)(64 * waves));
  const size_t shm = grid_batched_lds(a.d, a.T, a.taps, cg, a.rowstride);
  if (a.d == 2) {
    if (cg == 8) grid_batched_taps<8, 2>(a, g, b, shm, s);
    else if (cg == 16) grid_batched_taps<16, 2>(a, g, b, shm, s);
    else grid_batched_taps<32, 2>(a, g, b, shm, s);
  } else {
    if (cg == 8) grid_batched_taps<8, 3>(a, g, b, shm, s);
    else if (cg == 16) grid_batched_taps<16, 3>(a, g, b, shm, s);
    else grid_batched_taps<32, 3>(a, g, b, shm, s);
  }
}

size_t interp_lds_bytes(int T, int taps, int CG, int rowstride, int maxs) {
  const int E = T + taps - 1;
  return sizeof(float2) * (size_t)E * rowstride + sizeof(float) * (size_t)maxs * 2 * taps +
         sizeof(int) * (size_t)maxs * 2;
}

template <int CG, int TT>
static void interp_lds_tt(const SampArgs& a, const int4* items, int nitems, int T, int rowstride, int maxs,
                          int nt0, dim3 g, size_t shm, hipStream_t s) {
  switch (a.taps) {
    case 4: hipLaunchKernelGGL((k_interp_lds<CG, 4, TT>), g, dim3(256), shm, s, a, items, nitems, T, rowstride, maxs, nt0); break;
    case 6: hipLaunchKernelGGL((k_interp_lds<CG, 6, TT>), g, dim3(256), shm, s, a, items, nitems, T, rowstride, maxs, nt0); break;
    case 8: hipLaunchKernelGGL((k_interp_lds<CG, 8, TT>), g, dim3(256), shm, s, a, items, nitems, T, rowstride, maxs, nt0); break;
    case 10: hipLaunchKernelGGL((k_interp_lds<CG, 10, TT>), g, dim3(256), shm, s, a, items, nitems, T, rowstride, maxs, nt0); break;
    default: break;
  }
}

template <int CG>
static void interp_lds_taps(const SampArgs& a, const int4* items, int nitems, int T, int rowstride, int maxs,
                            int nt0, dim3 g, size_t shm, hipStream_t s) {
  if (T == 8) interp_lds_tt<CG, 8>(a, items, nitems, T, rowstride, maxs, nt0, g, shm, s);
  else interp_lds_tt<CG, 0>(a, items, nitems, T, rowstride, maxs, nt0, g, shm, s);
}

template <int CG, int TAPS, int TT, int CPL>
static void interp_cs_go(const SampArgs& a, const int4* items, int nitems, int T, int rowstride, int maxs,
                         int nt0, dim3 g, size_t shm, int nth, hipStream_t s) {
  if (shm > 65536) {   // more than 64 KB of dynamic LDS (32 coils) needs the kernel's limit raised (once)
    static const bool raised = [] {
      return hipFuncSetAttribute(reinterpret_cast<const void*>(&k_interp_cs<CG, TAPS, TT, 256, CPL>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024) == hipSuccess &&
             hipFuncSetAttribute(reinterpret_cast<const void*>(&k_interp_cs<CG, TAPS, TT, 512, CPL>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024) == hipSuccess;
    }();
    (void)raised;
  }
  if (nth == 512)
    hipLaunchKernelGGL((k_interp_cs<CG, TAPS, TT, 512, CPL>), g, dim3(512), shm, s, a, items, nitems, T, rowstride,
                       maxs, nt0);
  else
    hipLaunchKernelGGL((k_interp_cs<CG, TAPS, TT, 256, CPL>), g, dim3(256), shm, s, a, items, nitems, T, rowstride,
                       maxs, nt0);
}

template <int CG, int TAPS, int TT>
static void interp_cs_cpl(const SampArgs& a, const int4* items, int nitems, int T, int rowstride, int maxs,
                          int nt0, dim3 g, size_t shm, int nth, int cpl, hipStream_t s) {
  if (cpl == 2) interp_cs_go<CG, TAPS, TT, 2>(a, items, nitems, T, rowstride, maxs, nt0, g, shm, nth, s);
  else interp_cs_go<CG, TAPS, TT, 1>(a, items, nitems, T, rowstride, maxs, nt0, g, shm, nth, s);
}

template <int CG, int TT>
static void interp_cs_tt(const SampArgs& a, const int4* items, int nitems, int T, int rowstride, int maxs,
                         int nt0, dim3 g, size_t shm, int nth, int cpl, hipStream_t s) {
  switch (a.taps) {
    case 4: interp_cs_cpl<CG, 4, TT>(a, items, nitems, T, rowstride, maxs, nt0, g, shm, nth, cpl, s); break;
    case 6: interp_cs_cpl<CG, 6, TT>(a, items, nitems, T, rowstride, maxs, nt0, g, shm, nth, cpl, s); break;
    case 8: interp_cs_cpl<CG, 8, TT>(a, items, nitems, T, rowstride, maxs, nt0, g, shm, nth, cpl, s); break;
    case 10: interp_cs_cpl<CG, 10, TT>(a, items, nitems, T, rowstride, maxs, nt0, g, shm, nth, cpl, s); break;
    default: break;
  }
}

template <int CG>
static void interp_cs_taps(const SampArgs& a, const int4* items, int nitems, int T, int rowstride, int maxs,
                           int nt0, dim3 g, size_t shm, int nth, int cpl, hipStream_t s) {
  if (T == 8) interp_cs_tt<CG, 8>(a, items, nitems, T, rowstride, maxs, nt0, g, shm, nth, cpl, s);
  else interp_cs_tt<CG, 0>(a, items, nitems, T, rowstride, maxs, nt0, g, shm, nth, cpl, s);
}

// lanes: 0 = (coil, tap row slot) k_interp_lds; 1, 2 = (coil, sample) k_interp_cs in blocks of 256, 512;
// 3, 4 = (coil pair, sample) in blocks of 256, 512
void launch_interp_lds(const SampArgs& a, const int4* items, int nitems, int T, int nt0, int cg, int maxs,
                       int ncgroups, int lanes, hipStream_t s) {
  if (nitems == 0) return;
  const int rowstride = (T + a.taps - 1) * cg + ((cg == 8) ? 8 : 0);
  const size_t shm = interp_lds_bytes(T, a.taps, cg, rowstride, maxs);
  const dim3 g((unsigned)nitems, (unsigned)ncgroups);
  if (lanes >= 1) {   // the region alone in LDS
    const size_t rshm = sizeof(float2) * (size_t)(T + a.taps - 1) * rowstride;
    const int nth = (lanes == 2 || lanes == 4) ? 512 : 256;
    const int cpl = lanes >= 3 ? 2 : 1;
    if (cg == 8) interp_cs_taps<8>(a, items, nitems, T, rowstride, maxs, nt0, g, rshm, nth, cpl, s);
    else if (cg == 16) interp_cs_taps<16>(a, items, nitems, T, rowstride, maxs, nt0, g, rshm, nth, cpl, s);
    else interp_cs_taps<32>(a, items, nitems, T, rowstride, maxs, nt0, g, rshm, nth, cpl, s);
    return;
  }
  if (cg == 8) interp_lds_taps<8>(a, items, nitems, T, rowstride, maxs, nt0, g, shm, s);
  else if (cg == 16) interp_lds_taps<16>(a, items, nitems, T, rowstride, maxs, nt0, g, shm, s);
  else interp_lds_taps<32>(a, items, nitems, T, rowstride, maxs, nt0, g, shm, s);
}

// ---------------------------------------------------------------------------
// forward B in 3D with the tile's extended region ((T + taps - 1)^3 cells x CG
// coils) staged in LDS.  Work item = (frame, tile, samples whose floor(n x)
// lies in the tile); lanes = (coil, tap row (s1, s2)), a shuffle reduction
// over the rows.
template <int CG, int TAPS, int NT, int TT>
__global__ void __launch_bounds__(NT, NT / 128) k_interp_lds3(SampArgs a, const int4* items, int T_, int maxs,
                                                              int nt0, int nt1) {
  constexpr int RS = 64 / CG;
  extern __shared__ float2 sh[];
  const int4 it = items[blockIdx.x];            // frame, tile, sbeg, send
  if (a.fdone && a.fdone[it.x]) return;
  // TT > 0: the tile side as a constant (the region width E too), so that the
  // tap-row addressing and the staging's index splits are shifts and adds
  const int T = TT > 0 ? TT : T_;
  const int E = T + TAPS - 1;
  const int m = TAPS / 2 - 1;
  const int E3 = E * E * E;
  float2* ext = sh;                                     // [E][E][E][CG]
  float* wpsi = (float*)(sh + (size_t)E3 * CG);         // [maxs][3*TAPS]
  int* wrel = (int*)(wpsi + (size_t)maxs * 3 * TAPS);   // [maxs][3]
  const int cbase = blockIdx.y * CG;
  const int t0 = it.y % nt0, t1 = (it.y / nt0) % nt1, t2 = it.y / (nt0 * nt1);
  const int org0 = t0 * T - m, org1 = t1 * T - m, org2 = t2 * T - m;
  const int64_t gbase = (int64_t)it.x * a.ntot;
  {
    constexpr int U = 8, STEP = NT / CG;
    const int c = threadIdx.x % CG;
    const bool cok = cbase + c < a.C;
    for (int e00 = threadIdx.x / CG; e00 < E3; e00 += STEP * U) {
      float2 v[U];
      int dst[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int e = e00 + u * STEP;
        v[u] = make_float2(0.f, 0.f);
        dst[u] = -1;
        if (e < E3) {
          const int e2 = e / (E * E), r = e - e2 * E * E, e1 = r / E, e0 = r - e1 * E;
          int l0 = org0 + e0, l1 = org1 + e1, l2 = org2 + e2;   // one wrap suffices
          if (l0 < 0) l0 += a.n0; else if (l0 >= a.n0) l0 -= a.n0;
          if (l1 < 0) l1 += a.n1; else if (l1 >= a.n1) l1 -= a.n1;
          if (l2 < 0) l2 += a.n2; else if (l2 >= a.n2) l2 -= a.n2;
          if (cok) v[u] = a.grid[(gbase + l0 + (int64_t)a.n0 * (l1 + (int64_t)a.n1 * l2)) * a.C + cbase + c];
          dst[u] = e * CG + c;
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (dst[u] >= 0) ext[dst[u]] = v[u];
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int cl = lane % CG, rs = lane / CG;
  const int c = cbase + cl;
  // the item's samples in batches of maxs: window values and offsets staged
  // per batch, the grid region once for the whole item
  for (int sb = it.z; sb < it.w; sb += maxs) {
  const int ns = min(maxs, it.w - sb);
  if (sb > it.z) lds_barrier();   // the previous batch is consumed
  for (int idx = threadIdx.x; idx < ns * 3 * TAPS; idx += blockDim.x)
    wpsi[idx] = a.psi[(int64_t)sb * 3 * TAPS + idx];
  for (int k = threadIdx.x; k < ns; k += blockDim.x) {
    const int smp = sb + k;
    int r0 = a.base[3 * smp] - org0, r1 = a.base[3 * smp + 1] - org1, r2 = a.base[3 * smp + 2] - org2;
    if (r0 < 0) r0 += a.n0; else if (r0 >= a.n0) r0 -= a.n0;
    if (r1 < 0) r1 += a.n1; else if (r1 >= a.n1) r1 -= a.n1;
    if (r2 < 0) r2 += a.n2; else if (r2 >= a.n2) r2 -= a.n2;
    wrel[3 * k] = r0;
    wrel[3 * k + 1] = r1;
    wrel[3 * k + 2] = r2;
  }
  lds_barrier();
  for (int k = wave; k < ns; k += NT / 64) {
    const float* ps = wpsi + k * 3 * TAPS;
    const int r0 = wrel[3 * k], r1 = wrel[3 * k + 1], r2 = wrel[3 * k + 2];
    float2 acc = make_float2(0.f, 0.f);
    for (int row = rs; row < TAPS * TAPS; row += RS) {
      const int s1 = row % TAPS, s2 = row / TAPS;
      const float2* rp = ext + ((size_t)((r2 + s2) * E + (r1 + s1)) * E + r0) * CG + cl;
      float2 racc = make_float2(0.f, 0.f);
#pragma unroll
      for (int s0 = 0; s0 < TAPS; ++s0) {
        const float2 v = rp[s0 * CG];
        const float p0 = ps[s0];
        racc.x += p0 * v.x;
        racc.y += p0 * v.y;
      }
      const float wr = ps[TAPS + s1] * ps[2 * TAPS + s2];
      acc.x += wr * racc.x;
      acc.y += wr * racc.y;
    }
    if (CG == 2) {
      // within each 16-lane row: lanes l and l ^ 2 (quad permutation), then
      // l + 4 and l + 8 (row shifts) -- lanes 0 and 1 of a row hold its sums
      // for coils 0 and 1; then across the four rows by row swaps
      acc = dpp_add2<0x4E>(acc);
      acc = dpp_add2<0x104>(acc);
      acc = dpp_add2<0x108>(acc);
      acc = make_float2(swap_sum<16>(acc.x), swap_sum<16>(acc.y));
      acc = make_float2(swap_sum<32>(acc.x), swap_sum<32>(acc.y));
    } else {
      acc = make_float2(class_sum<CG>(acc.x), class_sum<CG>(acc.y));
    }
    if (rs == 0 && c < a.C) a.y[(int64_t)(sb + k) * a.C + c] = acc;
  }
  }
}

// forward B in 3D with the region of 2 coils staged as in k_interp_lds3 (one
// 16-byte cell = both coils) and lanes over samples and tap planes.
// SPLIT = 0: a wave takes 8 samples, lane (s, z) sums the TAPS^2 taps of plane
// z of sample s for both coils; SPLIT = 1: a wave takes 4 samples, lane
// (s, z, h) the taps of plane z with s0 = h mod 2, and the region's planes are
// PS = E^2 + 1 cells apart, so that the 16 lanes of a sample read 16 distinct
// 4-bank groups (all 64 banks) at every step.  The lanes of a sample are summed
// by DPP steps (three or four, in place of five over 32 tap-row lanes).  Window
// values and base cells come from global memory, one address per sample
// broadcast to its lanes, so only the region sits in LDS.
template <int TAPS, int NT, int TT, int SPLIT>
__global__ void __launch_bounds__(NT, NT / 128) k_interp_cs3(SampArgs a, const int4* items, int T_, int nt0,
                                                             int nt1, int ppad) {
  static_assert(TAPS == 8, "8 planes per sample: one per lane group");
  constexpr int CG = 2, NW = NT / 64, LPS = SPLIT ? 16 : 8, SPW = 64 / LPS;
  extern __shared__ float2 sh[];
  const int4 it = items[blockIdx.x];            // frame, tile, sbeg, send
  if (a.fdone && a.fdone[it.x]) return;
  const int T = TT > 0 ? TT : T_;
  const int E = T + TAPS - 1;
  const int PS = E * E + (SPLIT ? ppad : 0);     // plane stride in cells
  const int m = TAPS / 2 - 1;
  const int E3 = E * E * E;
  float2* ext = sh;                                     // [E][PS][2]
  const int cbase = blockIdx.y * CG;
  const int t0 = it.y % nt0, t1 = (it.y / nt0) % nt1, t2 = it.y / (nt0 * nt1);
  const int org0 = t0 * T - m, org1 = t1 * T - m, org2 = t2 * T - m;
  const int64_t gbase = (int64_t)it.x * a.ntot;
  {
    constexpr int U = 8, STEP = NT / CG;
    const int c = threadIdx.x % CG;
    const bool cok = cbase + c < a.C;
    for (int e00 = threadIdx.x / CG; e00 < E3; e00 += STEP * U) {
      float2 v[U];
      int dst[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int e = e00 + u * STEP;
        v[u] = make_float2(0.f, 0.f);
        dst[u] = -1;
        if (e < E3) {
          const int e2 = e / (E * E), r = e - e2 * E * E, e1 = r / E, e0 = r - e1 * E;
          int l0 = org0 + e0, l1 = org1 + e1, l2 = org2 + e2;   // one wrap suffices
          if (l0 < 0) l0 += a.n0; else if (l0 >= a.n0) l0 -= a.n0;
          if (l1 < 0) l1 += a.n1; else if (l1 >= a.n1) l1 -= a.n1;
          if (l2 < 0) l2 += a.n2; else if (l2 >= a.n2) l2 -= a.n2;
          if (cok) v[u] = a.grid[(gbase + l0 + (int64_t)a.n0 * (l1 + (int64_t)a.n1 * l2)) * a.C + cbase + c];
          dst[u] = (e2 * PS + r) * CG + c;
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (dst[u] >= 0) ext[dst[u]] = v[u];
    }
  }
  lds_barrier();
  const float4* ext4 = reinterpret_cast<const float4*>(ext);   // cell -> (coil 0, coil 1)
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int h = SPLIT ? (lane & 1) : 0;
  const int pl = SPLIT ? ((lane >> 1) & 7) : (lane & 7);
  const int sl = lane / LPS;
  const int ns = it.w - it.z;
  for (int k0 = wave * SPW; k0 < ns; k0 += NW * SPW) {
    const int k = k0 + sl;
    const int smp = it.z + (k < ns ? k : 0);   // lanes past the item's end repeat its first sample
    const float* ps = a.psi + (int64_t)smp * 3 * TAPS;
    constexpr int NX = SPLIT ? TAPS / 2 : TAPS;
    float px[NX], py[TAPS];
#pragma unroll
    for (int j = 0; j < NX; ++j) px[j] = ps[SPLIT ? 2 * j + h : j];
#pragma unroll
    for (int s1 = 0; s1 < TAPS; ++s1) py[s1] = ps[TAPS + s1];
    const float pz = ps[2 * TAPS + pl];
    int r0 = a.base[3 * smp] - org0, r1 = a.base[3 * smp + 1] - org1, r2 = a.base[3 * smp + 2] - org2;
    if (r0 < 0) r0 += a.n0; else if (r0 >= a.n0) r0 -= a.n0;
    if (r1 < 0) r1 += a.n1; else if (r1 >= a.n1) r1 -= a.n1;
    if (r2 < 0) r2 += a.n2; else if (r2 >= a.n2) r2 -= a.n2;
    const float4* pp = ext4 + (r2 + pl) * PS + r1 * E + r0 + h;
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int s1 = 0; s1 < TAPS; ++s1) {
      float4 racc = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
      for (int j = 0; j < NX; ++j) {
        const float4 v = pp[s1 * E + (SPLIT ? 2 * j : j)];
        racc.x += px[j] * v.x;
        racc.y += px[j] * v.y;
        racc.z += px[j] * v.z;
        racc.w += px[j] * v.w;
      }
      acc.x += py[s1] * racc.x;
      acc.y += py[s1] * racc.y;
      acc.z += py[s1] * racc.z;
      acc.w += py[s1] * racc.w;
    }
    float2 a0 = make_float2(pz * acc.x, pz * acc.y), a1 = make_float2(pz * acc.z, pz * acc.w);
    // over the sample's lanes: l ^ 1, l ^ 2 (quad permutations), l + 4 (row
    // shift); SPLIT: also l + 8 -- the sample's first lane holds the sums
    a0 = dpp_add2<0xB1>(a0);
    a1 = dpp_add2<0xB1>(a1);
    a0 = dpp_add2<0x4E>(a0);
    a1 = dpp_add2<0x4E>(a1);
    a0 = dpp_add2<0x104>(a0);
    a1 = dpp_add2<0x104>(a1);
    if (SPLIT) {
      a0 = dpp_add2<0x108>(a0);
      a1 = dpp_add2<0x108>(a1);
    }
    if ((lane % LPS) == 0 && k < ns) {
      float2* dst = a.y + (int64_t)smp * a.C + cbase;
      dst[0] = a0;
      if (cbase + 1 < a.C) dst[1] = a1;
    }
  }
}

size_t interp_lds3_bytes(int T, int taps, int cg, int maxs) {
  const size_t E = (size_t)T + taps - 1;
  return sizeof(float2) * E * E * E * cg + sizeof(float) * (size_t)maxs * 3 * taps + sizeof(int) * (size_t)maxs * 3;
}

template <int NT>
static void interp_lds3_launch(const SampArgs& a, const int4* items, dim3 g, size_t shm, int T, int nt0, int nt1,
                               int cg, int maxs, hipStream_t s) {
#define CGS_IL3(CG_, TP)                                                                                  \
  do {                                                                                                  \
    if (T == 8) hipLaunchKernelGGL((k_interp_lds3<CG_, TP, NT, 8>), g, dim3(NT), shm, s, a, items, T, maxs, nt0, nt1); \
    else hipLaunchKernelGGL((k_interp_lds3<CG_, TP, NT, 0>), g, dim3(NT), shm, s, a, items, T, maxs, nt0, nt1);        \
  } while (0)
  if (cg == 2) {
    switch (a.taps) { case 4: CGS_IL3(2, 4); break; case 6: CGS_IL3(2, 6); break;
                      case 8: CGS_IL3(2, 8); break; case 10: CGS_IL3(2, 10); break; default: break; }
  } else {
    switch (a.taps) { case 4: CGS_IL3(4, 4); break; case 6: CGS_IL3(4, 6); break;
                      case 8: CGS_IL3(4, 8); break; case 10: CGS_IL3(4, 10); break; default: break; }
  }
#undef CGS_IL3
}

// threads per block: the staged region (54 KB at taps 8) allows two blocks per
// CU, so 1024-thread blocks give 8 waves per SIMD to cover the LDS latency of
// the tap reads (256-thread blocks: 2 waves per SIMD, 1.34x slower on the
// 128^3 koosh ball)
// the region alone in LDS (54 KB at taps 8: two blocks per CU); split: the
// 16-lanes-per-sample layout with padded planes
void launch_interp_cs3(const SampArgs& a, const int4* items, int nitems, int T, int nt0, int nt1, int threads,
                       int split, hipStream_t s) {
  if (nitems == 0 || a.taps != 8) return;
  // cells of padding after each E^2-cell plane (split layout; CGS_INTERP3_PAD, default 1)
  static const int ppad = [] { const char* e = std::getenv("CGS_INTERP3_PAD"); return e ? std::max(0, std::min(15, std::atoi(e))) : 1; }();
  const size_t E = (size_t)T + a.taps - 1;
  const size_t shm = sizeof(float2) * E * (E * E + (split ? ppad : 0)) * 2;
  const dim3 g((unsigned)nitems, (unsigned)((a.C + 1) / 2));
#define CGS_IC3(NT_, SP_)                                                                                        \
  do {                                                                                                         \
    if (T == 8) hipLaunchKernelGGL((k_interp_cs3<8, NT_, 8, SP_>), g, dim3(NT_), shm, s, a, items, T, nt0, nt1, ppad); \
    else hipLaunchKernelGGL((k_interp_cs3<8, NT_, 0, SP_>), g, dim3(NT_), shm, s, a, items, T, nt0, nt1, ppad);        \
  } while (0)
  if (split) {
    if (threads >= 1024) CGS_IC3(1024, 1);
    else if (threads >= 512) CGS_IC3(512, 1);
    else CGS_IC3(256, 1);
  } else {
    if (threads >= 1024) CGS_IC3(1024, 0);
    else if (threads >= 512) CGS_IC3(512, 0);
    else CGS_IC3(256, 0);
  }
#undef CGS_IC3
}

void launch_interp_lds3(const SampArgs& a, const int4* items, int nitems, int T, int nt0, int nt1, int cg, int maxs,
                        int threads, hipStream_t s) {
  if (nitems == 0) return;
  const size_t shm = interp_lds3_bytes(T, a.taps, cg, maxs);
  const dim3 g((unsigned)nitems, (unsigned)((a.C + cg - 1) / cg));
  if (threads >= 1024) interp_lds3_launch<1024>(a, items, g, shm, T, nt0, nt1, cg, maxs, s);
  else if (threads >= 512) interp_lds3_launch<512>(a, items, g, shm, T, nt0, nt1, cg, maxs, s);
  else interp_lds3_launch<256>(a, items, g, shm, T, nt0, nt1, cg, maxs, s);
}

// ---------------------------------------------------------------------------
template <int CG, int D>
static void interp_taps(const SampArgs& a, dim3 g, hipStream_t s) {
  switch (a.taps) {
    case 4: hipLaunchKernelGGL((k_interp<CG, 4, D>), g, dim3(256), 0, s, a); break;
    case 6: hipLaunchKernelGGL((k_interp<CG, 6, D>), g, dim3(256), 0, s, a); break;
    case 8: hipLaunchKernelGGL((k_interp<CG, 8, D>), g, dim3(256), 0, s, a); break;
    case 10: hipLaunchKernelGGL((k_interp<CG, 10, D>), g, dim3(256), 0, s, a); break;
    default: break;
  }
}

void launch_interp(const SampArgs& a, int cg, hipStream_t s) {
  const dim3 g((unsigned)((a.nitems + 3) / 4));
  if (a.nitems == 0) return;
  if (a.d == 2) {
    if (cg == 8) interp_taps<8, 2>(a, g, s);
    else if (cg == 16) interp_taps<16, 2>(a, g, s);
    else interp_taps<32, 2>(a, g, s);
  } else {
    if (cg == 8) interp_taps<8, 3>(a, g, s);
    else if (cg == 16) interp_taps<16, 3>(a, g, s);
    else interp_taps<32, 3>(a, g, s);
  }
}

template <int CG, int D>
static void grid_taps(const GridArgs& a, dim3 g, dim3 b, size_t shm, hipStream_t s) {
  switch (a.taps) {
    case 4: hipLaunchKernelGGL((k_grid<CG, 4, D>), g, b, shm, s, a); break;
    case 6: hipLaunchKernelGGL((k_grid<CG, 6, D>), g, b, shm, s, a); break;
    case 8: hipLaunchKernelGGL((k_grid<CG, 8, D>), g, b, shm, s, a); break;
    case 10: hipLaunchKernelGGL((k_grid<CG, 10, D>), g, b, shm, s, a); break;
    default: break;
  }
}

void launch_grid(const GridArgs& a, int cg, int waves, int nitems, int ncgroups, hipStream_t s) {
  if (nitems == 0) return;
  const dim3 g((unsigned)nitems, (unsigned)ncgroups), b((unsigned)(64 * waves));
  const int nrows_tile = (a.d == 3) ? a.T * a.T : a.T;
  const size_t shm = sizeof(float2) * (size_t)nrows_tile * a.rowstride;
  if (a.d == 2) {
    if (cg == 8) grid_taps<8, 2>(a, g, b, shm, s);
    else if (cg == 16) grid_taps<16, 2>(a, g, b, shm, s);
    else grid_taps<32, 2>(a, g, b, shm, s);
  } else {
    if (cg == 8) grid_taps<8, 3>(a, g, b, shm, s);
    else if (cg == 16) grid_taps<16, 3>(a, g, b, shm, s);
    else grid_taps<32, 3>(a, g, b, shm, s);
  }
}

void launch_grid_reduce(const Plan& p, const float2* partial, float2* grid, const int* fl, hipStream_t s) {
  if (p.n_reduce_items == 0) return;
  const int64_t tilecells = (int64_t)p.T * p.T * (p.d == 3 ? p.T : 1);
  const int64_t outs = tilecells * p.C;
  const unsigned gy = (unsigned)std::max<int64_t>(1, std::min<int64_t>(64, (outs + 255) / 256));
  hipLaunchKernelGGL(k_grid_reduce, dim3((unsigned)p.n_reduce_items, gy), dim3(256), 0, s,
                     p.reduce_items, partial, grid, p.C, p.T, p.d, p.n[0], p.n[1], p.n[2],
                     p.nt[0], p.nt[1], p.ntot, tilecells, fl);
}

}  // namespace cgs
