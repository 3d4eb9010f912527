// kernels_sweep.hip -- 2D adjoint window scatter (B^T, the gridding of
// nfft_adjoint, NFFTTraits.hpp:201-204) as a sweep along grid lines.
//
// One wave owns one segment of one grid line: the W cells [S, S + W) of line
// l1 (axis 1 fixed, cells along axis 0, which are adjacent in memory).  Its
// lanes are the float components (re, im of every coil) of one cell, so every
// FMA a lane issues is useful work.  The samples whose footprint covers the line
// and meets the segment ("entries") are presorted on the host by their first
// footprint cell r0 relative to S.  The wave sweeps r0 upwards with a window of
// TAPS accumulator registers: the window slot of cell c is c mod TAPS, and once
// the entries starting at cell c are applied, cell c is complete and is stored
// (each grid cell is written exactly once, by one wave, in a fixed order: no
// atomics, no zero-fill of the grid, deterministic).
//
// The axis-0 window values of a sample are stored rotated (psi_rot[s][k] = the
// value of the tap that lands on a cell == k mod TAPS), so an entry updates
// slot k with psi_rot[s][k] whatever its start cell: the register index of every
// FMA is static.  This needs n0 and W to be multiples of TAPS.
//
// With fewer than 32 coils the wave's 64 lanes hold TS = 64 / GL groups of GL
// float components (GL = 2 x coils per wave, rounded up to a power of two);
// group g owns window slots [g NSL, (g + 1) NSL), NSL = TAPS / TS.
#include <cstdlib>

#include "cgs_internal.hpp"

namespace cgs {

// Entries are taken P at a time: the chunk's loads (records, window values,
// data) are all issued before its first FMA, so a wave keeps P loads in
// flight.  Before an entry that starts at cell r0 every cell below r0 is
// complete and is flushed; the flush picks its (static) register through a
// uniform branch on the slot.
template <int TAPS, int TS>
__device__ __forceinline__ void sweep_flush(float (&acc)[TAPS / TS], int c, int ca, int cb, int g, bool live,
                                            float* out, int ncomp) {
  constexpr int NSL = TAPS / TS;
  const int slot = ((c % TAPS) + TAPS) % TAPS;
#pragma unroll
  for (int k = 0; k < TAPS; ++k) {
    if (k == slot) {
      if (g == k / NSL) {
        if (c >= ca && c < cb && live) out[(int64_t)c * ncomp] = acc[k % NSL];
        acc[k % NSL] = 0.f;
      }
    }
  }
}

template <int TAPS, int TS, bool WEIGHTED, int P>
__global__ void __launch_bounds__(256) k_grid_sweep(SweepArgs a) {
  constexpr int NSL = TAPS / TS;   // window slots of one lane group
  constexpr int GL = 64 / TS;      // lanes (float components) per group
  const int lane = threadIdx.x & 63;
  // wave-uniform (the compiler cannot prove it of threadIdx.x >> 6): scalar
  // registers, scalar loads and scalar branches for everything derived from it
  const int item = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
  if (item >= a.nitems) return;    // the block's waves never synchronise
  const int4 it = as_const(a.items)[2 * item];       // {entry begin, entry end, frame, line * nseg + segment}
  const int4 it2 = as_const(a.items)[2 * item + 1];  // {partial offset or -1, first cell, end cell, 0}
  if (it.z < 0) return;                              // padding slot
  if (a.fdone && a.fdone[it.z]) return;
  const int g = lane / GL, q = lane - (lane / GL) * GL;
  const int comp = blockIdx.y * GL + q;   // float component of a sample / cell
  const int ncomp = 2 * a.C;
  const bool live = comp < ncomp;
  const int l1 = it.w / a.nseg;
  const int S = (it.w - l1 * a.nseg) * a.W;
  const int E = min(a.W, a.n0 - S);
  const float* yf = reinterpret_cast<const float*>(a.y) + (live ? comp : 0);
  // cells [ca, cb) of the segment go to the grid, or (a piece of a split
  // segment) to its partial range
  const int ca = it2.y, cb = it2.z;
  float* out = it2.x < 0
                   ? reinterpret_cast<float*>(a.grid + ((int64_t)it.z * a.ntot + (int64_t)l1 * a.n0 + S) * a.C) + comp
                   : a.part + it2.x - (int64_t)ca * ncomp + comp;
  float acc[NSL];
#pragma unroll
  for (int j = 0; j < NSL; ++j) acc[j] = 0.f;
  const const_ptr<int4> ent = as_const(a.ent);
  int c = -(TAPS - 1);   // every cell below c is flushed
  // Software pipeline over chunks of P entries: the data of chunk k + 1 and
  // the records of chunk k + 2 are in flight while chunk k is applied.  The
  // entry table ends with kSweepPad >= 3 P padding records (sample 0), so
  // records past this item's end are valid addresses: loaded and skipped.
  int4 rn[P], rnn[P];
  float yn[P];
#pragma unroll
  for (int j = 0; j < P; ++j) rn[j] = ent[it.x + j];
#pragma unroll
  for (int j = 0; j < P; ++j) yn[j] = yf[(int64_t)rn[j].x * ncomp];
#pragma unroll
  for (int j = 0; j < P; ++j) rnn[j] = ent[it.x + P + j];
  for (int e = it.x; e < it.y; e += P) {
    const int n = min(P, it.y - e);
    int4 r[P];
    float yv[P];
    float pr[P][NSL];
#pragma unroll
    for (int j = 0; j < P; ++j) {
      r[j] = rn[j];
      yv[j] = yn[j];
      rn[j] = rnn[j];
    }
#pragma unroll
    for (int j = 0; j < P; ++j) yn[j] = yf[(int64_t)rn[j].x * ncomp];
#pragma unroll
    for (int j = 0; j < P; ++j) rnn[j] = ent[e + 2 * P + j];
    // window values through the scalar cache (a lane-uniform vector load
    // would return 64 copies); lane group g takes taps g NSL .. g NSL + NSL - 1
#pragma unroll
    for (int j = 0; j < P; ++j) {
      const const_ptr<float> pc = as_const(a.psi_rot) + (int64_t)r[j].x * TAPS;
      float ps[TAPS];
#pragma unroll
      for (int k = 0; k < TAPS; ++k) ps[k] = pc[k];
#pragma unroll
      for (int k = 0; k < NSL; ++k) {
        float v = ps[k];
#pragma unroll
        for (int gg = 1; gg < TS; ++gg) v = (g == gg) ? ps[gg * NSL + k] : v;
        pr[j][k] = v;
      }
    }
#pragma unroll
    for (int j = 0; j < P; ++j) {
      if (j < n) {
        for (; c < r[j].y; ++c) sweep_flush<TAPS, TS>(acc, c, ca, cb, g, live, out, ncomp);
        const float yy = __int_as_float(WEIGHTED ? r[j].w : r[j].z) * yv[j];
#pragma unroll
        for (int k = 0; k < NSL; ++k) acc[k] = fmaf(pr[j][k], yy, acc[k]);
      }
    }
  }
  for (; c < cb; ++c) sweep_flush<TAPS, TS>(acc, c, ca, cb, g, live, out, ncomp);
}

// the cells of split line segments: the sum of the pieces' partial ranges that
// cover each cell, in piece order (0 where none does); lanes = float components
__global__ void __launch_bounds__(256) k_grid_sweep_reduce(SweepArgs a) {
  const int item = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
  if (item >= a.nritems) return;
  const int4 it = as_const(a.ritems)[item];   // {piece begin, piece end, frame, line * nseg + segment}
  if (a.fdone && a.fdone[it.z]) return;
  const int comp = blockIdx.y * 64 + (threadIdx.x & 63);
  const int ncomp = 2 * a.C;
  if (comp >= ncomp) return;
  const int l1 = it.w / a.nseg;
  const int S = (it.w - l1 * a.nseg) * a.W;
  const int E = min(a.W, a.n0 - S);
  float* out = reinterpret_cast<float*>(a.grid + ((int64_t)it.z * a.ntot + (int64_t)l1 * a.n0 + S) * a.C) + comp;
  const const_ptr<int4> pcs = as_const(a.rpieces);
  int j0 = it.x;   // first piece whose range may still cover the current cell
  for (int c = 0; c < E; ++c) {
    while (j0 < it.y && pcs[j0].z <= c) ++j0;
    float acc = 0.f;
    for (int j = j0; j < it.y; ++j) {
      const int4 pc = pcs[j];   // {offset, first cell, end cell, 0}; first cells do not decrease
      if (pc.y > c) break;
      if (c < pc.z) acc += a.part[pc.x + (int64_t)(c - pc.y) * ncomp + comp];
    }
    out[(int64_t)c * ncomp] = acc;
  }
}

// entries per chunk (CGS_SWEEP_P: 4 or 6; scalar registers hold the chunk's
// records, the next chunk's and the window values)
static int sweep_chunk() {
  static const int p = [] {
    const char* e = std::getenv("CGS_SWEEP_P");
    const int v = e ? std::atoi(e) : 4;
    return v == 6 ? 6 : 4;
  }();
  return p;
}

template <int TAPS, int TS>
static void sweep_launch(const SweepArgs& a, dim3 g, hipStream_t s) {
  const bool p4 = sweep_chunk() == 4;
  if (a.weighted) {
    if (p4) hipLaunchKernelGGL((k_grid_sweep<TAPS, TS, true, 4>), g, dim3(256), 0, s, a);
    else hipLaunchKernelGGL((k_grid_sweep<TAPS, TS, true, 6>), g, dim3(256), 0, s, a);
  } else {
    if (p4) hipLaunchKernelGGL((k_grid_sweep<TAPS, TS, false, 4>), g, dim3(256), 0, s, a);
    else hipLaunchKernelGGL((k_grid_sweep<TAPS, TS, false, 6>), g, dim3(256), 0, s, a);
  }
}

template <int TAPS>
static bool sweep_ts(const SweepArgs& a, int ts, dim3 g, hipStream_t s) {
  switch (ts) {
    case 1: sweep_launch<TAPS, 1>(a, g, s); return true;
    case 2: if constexpr (TAPS % 2 == 0) { sweep_launch<TAPS, 2>(a, g, s); return true; } return false;
    case 4: if constexpr (TAPS % 4 == 0) { sweep_launch<TAPS, 4>(a, g, s); return true; } return false;
    case 8: if constexpr (TAPS % 8 == 0) { sweep_launch<TAPS, 8>(a, g, s); return true; } return false;
    default: return false;
  }
}

// lane groups per wave for C coils: GL = 2C rounded up to a power of two, at
// most 64 (more coils take several waves, blockIdx.y)
int sweep_groups(int C) {
  int gl = 2;
  while (gl < 2 * C && gl < 64) gl *= 2;
  return 64 / gl;
}

bool grid_sweep_supported(int taps, int C) {
  const int ts = sweep_groups(C);
  return (taps == 4 || taps == 6 || taps == 8 || taps == 10) && taps % ts == 0;
}

void launch_grid_sweep_reduce(const SweepArgs& a, hipStream_t s) {
  if (a.nritems == 0) return;
  const dim3 g((unsigned)((a.nritems + 3) / 4), (unsigned)((2 * a.C + 63) / 64));
  hipLaunchKernelGGL(k_grid_sweep_reduce, g, dim3(256), 0, s, a);
}

void launch_grid_sweep(const SweepArgs& a, hipStream_t s) {
  if (a.nitems == 0) return;
  const int ts = sweep_groups(a.C);
  const int gl = 64 / ts;
  const dim3 g((unsigned)((a.nitems + 3) / 4), (unsigned)((2 * a.C + gl - 1) / gl));
  switch (a.taps) {
    case 4: sweep_ts<4>(a, ts, g, s); break;
    case 6: sweep_ts<6>(a, ts, g, s); break;
    case 8: sweep_ts<8>(a, ts, g, s); break;
    case 10: sweep_ts<10>(a, ts, g, s); break;
    default: break;
  }
}

}  // namespace cgs
