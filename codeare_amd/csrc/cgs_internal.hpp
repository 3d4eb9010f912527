// cgs_internal.hpp -- plan object and kernel launch interfaces of libcgsense_hip.
//
// Device data layouts (all complex64 as float2):
//   maps      S[i*C + c]                 coil fastest ("coil-interleaved")
//   grid      G[(f*ntot + l)*C + c]      l = l0 + n0*(l1 + n1*l2), oversampled k-space
//   coil imgs X[(f*Ntot + i)*C + c]      per-coil images (E^H before the coil sum)
//   samples   Y[s*C + c]                 s = tile-sorted sample position (all frames)
//   images    x[f*Ntot + i]              i = i0 + N0*(i1 + N1*i2), codeare order
// Samples are sorted by (frame, grid tile of the base cell); base[s*d+a] holds
// (floor(n_a*x_a) - m) mod n_a and psi[(s*d+a)*taps+t] the window value of tap t.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/cgsense.h"

namespace cgs {

constexpr int kFftMaxElem = 8192;   // lines*n held in LDS by one FFT block
constexpr int kFftThreads = 256;

struct GridItem {       // one gridding work item: part of one tile's visit list
  int frame, tile, vbeg, vend;
  int slot;             // -1: write the tile to the grid; else partial slot
  int pad0, pad1, pad2;
};
struct ReduceItem {     // sums the partial slots of one heavy tile
  int frame, tile, slot0, nslots;
};
// Workgroup barrier for data exchanged through LDS.  __syncthreads() is a
// release / acquire fence on every address space, so it also waits for the
// wave's outstanding global loads and stores (vmcnt 0): a store issued before
// an exchange, or a load issued ahead for later use, would be waited for at
// every barrier.  Kernels whose waves pass nothing to each other through
// global memory use this one: fences on the local address space only
// (lgkmcnt 0), then s_barrier.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Lane sums without LDS traffic: DPP lane moves within 16-lane rows and the
// gfx950 row swaps (v_permlane16_swap / v_permlane32_swap) across rows.  All
// lanes of the wave take part.
template <int CTRL>   // v of the lane that DPP control CTRL selects (0 outside the row)
__device__ __forceinline__ float dpp_mov(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, true));
}

// v + v of lane l ^ 16 (SPAN 16) or l ^ 32 (SPAN 32): with both operands v,
// the two results of a row swap hold the lane's own value and its partner's
template <int SPAN>
__device__ __forceinline__ float swap_sum(float v) {
  const unsigned u = __float_as_uint(v);
  if (SPAN == 16) {
    const auto r = __builtin_amdgcn_permlane16_swap(u, u, false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
  } else {
    const auto r = __builtin_amdgcn_permlane32_swap(u, u, false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
  }
}

// FP64 forms: the two 32-bit halves move separately (exact)
template <int CTRL>
__device__ __forceinline__ double dpp_mov_d(double v) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)u, CTRL, 0xF, 0xF, true);
  const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), CTRL, 0xF, 0xF, true);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

template <int SPAN>
__device__ __forceinline__ double swap_sum_d(double v) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)u, hi = (unsigned)(u >> 32);
  unsigned l0, l1, h0, h1;
  if (SPAN == 16) {
    const auto rl = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
    const auto rh = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
    l0 = rl[0]; l1 = rl[1]; h0 = rh[0]; h1 = rh[1];
  } else {
    const auto rl = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
    const auto rh = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
    l0 = rl[0]; l1 = rl[1]; h0 = rh[0]; h1 = rh[1];
  }
  return __longlong_as_double((long long)(((unsigned long long)h0 << 32) | l0)) +
         __longlong_as_double((long long)(((unsigned long long)h1 << 32) | l1));
}

// sum over all 64 lanes of the wave (every lane ends with the same sum)
__device__ __forceinline__ double wave_sum_d(double v) {
  v += dpp_mov_d<0xB1>(v);
  v += dpp_mov_d<0x4E>(v);
  v += dpp_mov_d<0x141>(v);
  v += dpp_mov_d<0x140>(v);
  v = swap_sum_d<16>(v);
  return swap_sum_d<32>(v);
}

// sum over the 16 lanes of a row (every lane ends with the same sum): quad
// permutations [1,0,3,2] and [2,3,0,1], then the half-row and row mirrors
__device__ __forceinline__ float row16_sum(float v) {
  v += dpp_mov<0xB1>(v);
  v += dpp_mov<0x4E>(v);
  v += dpp_mov<0x141>(v);
  v += dpp_mov<0x140>(v);
  return v;
}

// sum over the aligned group of GS consecutive lanes (GS a power of two
// <= 64; every lane ends with its group's sum); only the group's own lanes
// are read, so groups that have left a loop may be inactive
template <int GS>
__device__ __forceinline__ float group_sum(float v) {
  if (GS >= 2) v += dpp_mov<0xB1>(v);
  if (GS >= 4) v += dpp_mov<0x4E>(v);
  if (GS >= 8) v += dpp_mov<0x141>(v);
  if (GS >= 16) v += dpp_mov<0x140>(v);
  if (GS >= 32) v = swap_sum<16>(v);
  if (GS >= 64) v = swap_sum<32>(v);
  return v;
}

// sum over the lanes l' with l' % CG == l % CG (CG a power of two <= 32):
// quad permutations for lane distances 1 and 2, row rotations by 4 and 8,
// then the row swaps; every lane ends with its class's sum
template <int CG>
__device__ __forceinline__ float class_sum(float v) {
  if (CG < 2) v += dpp_mov<0xB1>(v);
  if (CG < 4) v += dpp_mov<0x4E>(v);
  if (CG < 8) v += dpp_mov<0x124>(v);
  if (CG < 16) v += dpp_mov<0x128>(v);
  if (CG < 32) v = swap_sum<16>(v);
  return swap_sum<32>(v);
}

struct CgState {        // per-frame CG scalars, FP64
  double rn, bn, xn, pq, alpha, beta, rel;
  int it, done, nan, pad;
  // beta folded into the next pad pass (lean Toeplitz path): the CG tail
  // snapshots rn and it and sets bpend; the pad pass forms beta and the stop
  // test from the tail's ||r||^2 partials and commits rn, it, beta, rel, done
  double rn_snap;
  int it_snap, bpend;
};

constexpr int kRedBlocks = 512;   // partial sums per frame for image-domain reductions
constexpr int kCoilBlocks = 2048; // blocks per frame of the coil combine

struct FftArgs {
  const float2* in;
  float2* out;
  int n, nin, nout, lb, ldn, sign;
  int pad_in, crop_out;   // input in image-index order (zero-pad) / output cropped
  // Toeplitz convolution pass: forward FFT, multiply by the real kernel
  // spectrum That[(f*conv_cells + r)*n + pos] (r = inner index / C), inverse FFT
  const float* That;
  int64_t conv_cells;
  int64_t inner, outers, outer_per_frame;
  const float2* tw;     // W_n^p = exp(-2 pi i p / n)
  int nst, radix[8];
  uint32_t ns_magic[8]; // ceil(2^32 / Ns) of each stage (Ns = product of earlier radices)
  int lb_log2;          // lb = 1 << lb_log2
  int use_reg;          // register-resident kernel (n = 256 or 512, 16 lines per block)
  // fused operations of the first forward / last adjoint pass (axis 0)
  const float2* x;      // image frames (load mode 1)
  const float2* S;      // maps (load mode 1); frame f's maps start at S + f*Sfs
  int64_t Sfs;
  const float* dinv0;
  const float* dinv1;
  const float* dinv2;
  int C, d, N0, N1, mask_nh;
  int64_t Ntot;
  const int* fdone;
  // fused CG direction update in load mode 1/4: x = r + beta * pold, written to
  // pnew by the block that holds coil 0
  const float2* r;
  const float2* pold;
  float2* pnew;
  const CgState* cgst;
  // fused CG tail (register kernel only): the pad pass writes ||p_new||^2 per
  // (frame, row) from the block holding coil 0; the crop+combine pass writes
  // Re<pdot, partial coil sum> per (frame, coil group, row)
  double* pp_part;
  double* pq_part;
  const float2* pdot;
};

// the three FFT passes of the 2D Toeplitz operator on the lean kernels
// (kernels_toep.hip): pad (x or p = r + beta pold, times S_c, FFT along axis 0
// of the 2N0 x N1 grid), conv (FFT along axis 1, times That, IFFT, in place),
// crop (IFFT along axis 0, crop, sum of conj(S_c) X_c over 16 coils).  Grid
// layout G[((f*N1 + i1)*2N0 + l0)*C + c]; image positions are not shifted.
struct ToepArgs {
  const float2* in;       // crop: grid
  float2* out;            // pad: grid; conv: grid (in place); crop: partial images P[(f*G + g)*Ntot + i]
  const float2* x;        // pad input image (when pold is null)
  const float2* r;        // pad: p_new = r + beta pold, written to pnew by coil 0
  const float2* pold;
  float2* pnew;
  const CgState* cgst;
  const float2* S;        // frame f's maps start at S + f*Sfs
  int64_t Sfs;
  const float* That;      // conv: That[(f*2N0 + l0)*2N1 + bin]
  const float2* tw;       // exp(-2 pi i p / n), n = 2 N of the pass axis
  const int* fdone;
  double* pp_part;        // pad: ||p_new||^2 per (f, i1)
  double* pq_part;        // crop: Re<pdot, P_g> per (f, g, i1)
  const float2* pdot;
  // pad with the folded beta: ||r||^2 partials of the CG tail (nrpart per
  // frame), stop rule, residual history; the frame's first block writes cgw,
  // hist and fdone_w
  const double* rpart;
  int nrpart, tolkind, maxit, hcap;
  int row;                // pad/crop: one block per grid row over all coil groups (crop: one partial image)
  int ppf;                // half passes: planes (f, i2) per frame
  double tol;
  CgState* cgw;
  double* hist;
  int* fdone_w;
  int C, N0, N1, F;
  int64_t Ntot;
  // CG tail folded into the passes (2D row kernels): the crop forms
  // q = A p + lambda p and the per-row <p, q>, <r, q>, |q|^2 and ||r||^2 of
  // the stored residual R; the next pad forms alpha = R / <p, q>,
  // ||r_new||^2 = R - 2 alpha <r, q> + alpha^2 |q|^2, beta and the stop test,
  // and updates x and r of its row before p = r + beta p
  int tailfold;
  const double* frr;      // pad: the crop's ||r||^2 row partials
  double* rr_part;        // crop output, per (f, i1)
  float lam;
  float2* xs;             // pad: the solution, updated in place
  float2* rs;             // pad: the residual, updated in place
  const float2* qv;       // pad: q of the previous iteration
  const double* fpq;      // pad: the crop's per-row partials (nrpart per frame)
  const double* frq;
  const double* fqq;
  const float2* rdot;     // crop: r (for <r, q>)
  double* rq_part;        // crop outputs, per (f, i1)
  double* qq_part;
};

struct SampArgs {
  const int* base;
  const float* psi;
  const float* w;       // density weights (sorted), nullable
  int d, taps, C;
  int n0, n1, n2;
  int64_t ntot;
  const float2* grid;
  float2* y;
  const int4* items;    // interp: {frame, start, count, 0}
  int nitems;
  const int* fdone;
};

struct GridArgs {
  const int* base;
  const float* psi;
  const float* w;
  const float2* y;
  float2* grid;
  float2* partial;
  const GridItem* items;
  const int* visits;
  const int* fdone;
  int d, taps, C, T, CGW, rowstride;
  int n0, n1, n2, nt0, nt1, nt2;
  int64_t ntot, tilecells;
};

// Binned gridding (kernels_bin.hip): every sample once, in the bin (TB^d cells)
// of its base cell; a work item adds its samples' footprints into the bin's
// region (E = TB + taps - 1 cells per axis) in LDS and writes the region to its
// slot of `reg`; a second pass sums, for every grid cell, the region cells that
// map onto it.
struct BinArgs {
  const int* ids;         // sorted sample ids grouped by (frame, bin)
  const int4* items;      // {frame * nbins + bin, first id, end id, region slot}, longest first
  const int2* range;      // per (frame, bin): {first slot, number of slots} (a bin's slots are consecutive)
  const float* brec;      // per id: its footprint offsets in the region (bits: r0 | r1 << 8 | r2 << 16)
                          // and its d x taps window values
  const float* w;         // density weights (of the sorted samples) or null
  const float2* y;        // Y[s * C + c]
  float2* reg;            // [item][group][E^d][CG]
  float2* grid;
  const int* fdone;
  int d, taps, C, CG, ngroups, TB, E, P1;
  int n0, n1, n2, nb0, nb1, nb2, F, nitems;
  int64_t ntot, ecells, nbins;
};

constexpr int kSweepPad = 24;   // padding records after the line-sweep entries (>= 3 x the chunk size)

// 2D gridding as a sweep along grid lines (kernels_sweep.hip): one wave per
// (frame, line l1, segment of W cells along axis 0)
struct SweepArgs {
  const int4* items;      // two int4 per item: {entry begin, entry end, frame (-1: padding),
                          // l1 * nseg + segment}, {partial offset in floats (-1: the grid),
                          // first cell, end cell, 0}
  const int4* ent;        // {sorted sample, r0, bits(phi1[t1]), bits(phi1[t1] * w)}, by r0 per item
  const float* psi_rot;   // [s * taps + k]: axis-0 window value of the tap on a cell == k mod taps
  const float2* y;        // Y[s * C + c]
  float2* grid;
  float* part;            // cell ranges of the pieces of split line segments
  const int4* ritems;     // split segments: {piece begin, piece end, frame, l1 * nseg + segment}
  const int4* rpieces;    // their pieces in order: {partial offset, first cell, end cell, 0}
  const int* fdone;
  int nitems, nritems, taps, C, W, nseg, n0, weighted;
  int64_t ntot;
};

struct FftTab {          // tables of one FFT length
  int n = 0;
  float2* tw = nullptr;   // exp(-2 pi i p / n)
  int radix[8] = {0};
  int nst = 0;
};

struct Plan {
  // geometry
  int d = 0, N[3] = {1, 1, 1}, n[3] = {1, 1, 1}, m = 0, taps = 0;
  int C = 0, F = 0;
  int64_t Mtot = 0, Ntot = 0, ntot = 0;
  unsigned flags = 0;
  double b[3] = {0, 0, 0};
  double sigma = 1.0;
  std::vector<int64_t> foff;          // F+1 sample offsets
  int T = 8, nt[3] = {1, 1, 1};       // tile side and tiles per axis
  int64_t tiles_per_frame = 1;
  int grid_cg = 8, grid_w = 4;        // gridding coil group / waves per block
  int interp_cg = 32;
  int visits_per_item = 256;
  bool use_interp_lds = true;         // gather from the tile staged in LDS
  int interp_lanes = 0;               // LDS gather lanes: 0 (coil, tap row); 2D (coil, sample) in blocks of
                                      // 1: 256, 2: 512 threads, (coil pair, sample) 3: 256, 4: 512;
                                      // 3D 1: (sample, plane), 2: (sample, plane, parity)
  bool use_grid_batched = true;       // batched visit lists (else one visit at a time)
  bool use_grid_reg = true;           // 2D: register accumulation, lane = cell of an 8x8 tile
  bool use_grid_mfma = false;         // 2D: scatter on the matrix cores (k_grid_mfma)
  int grid_cgr = 16;                  // coils per wave in the register kernel
  bool use_grid_atomic = false;       // LDS float atomics variant (kept for comparison)
  bool use_grid_sweep = false;        // 2D: line-sweep scatter (kernels_sweep.hip) where supported
  bool sweep_allowed = false;         // build the line-sweep tables (CGS_GRID_SWEEP=1)
  int sweep_W = 64;                   // cells per line segment (a multiple of taps)
  int tile_item_max = 64;             // samples per tile-level interp item
  int interp3_threads = 1024;         // 3D LDS gather: threads per block

  // stream / state
  hipStream_t stream = nullptr;
  bool own_stream = false;
  bool have_k = false, have_w = false, have_S = false;

  // constant tables
  FftTab fa[3];                       // FFT lengths n_a of the NFFT grid
  FftTab ft[3];                       // FFT lengths 2 N_a of the Toeplitz embedding
  float* dinv[3] = {nullptr, nullptr, nullptr};
  int64_t* d_foff = nullptr;

  // maps
  float2* S = nullptr;
  float* Imap = nullptr;
  int64_t Sfs = 0;                    // maps frame stride (Ntot*C with per-frame maps, else 0)
  int64_t Ifs = 0;                    // intensity-map frame stride (Ntot or 0)

  // samples (sorted)
  int* base = nullptr;
  float* psi = nullptr;
  float* w = nullptr;
  int* perm = nullptr;                // sorted -> original sample index
  int* iperm = nullptr;               // original -> sorted sample index
  int* sframe = nullptr;              // frame of each sorted sample
  std::vector<int> h_perm;
  int4* interp_items = nullptr;
  int n_interp_items = 0;
  int4* tile_items = nullptr;         // {frame, tile, sbeg, send}
  int n_tile_items = 0;
  GridItem* grid_items = nullptr;
  int n_grid_items = 0;
  ReduceItem* reduce_items = nullptr;
  int n_reduce_items = 0;
  int* visits = nullptr;
  int4* vrec = nullptr;               // per-visit records for the atomic gridding kernel
  int64_t n_visits = 0;
  int n_slots = 0;
  float2* partial = nullptr;
  // line-sweep scatter tables (2D)
  int4* sw_items = nullptr;           // two int4 per item (SweepArgs::items, items2)
  int n_sw_items = 0, sw_nseg = 0;
  int4* sw_ritems = nullptr;          // split line segments and their pieces
  int4* sw_rpieces = nullptr;
  int n_sw_ritems = 0;
  float* sw_part = nullptr;           // partial cell ranges of the pieces
  int sweep_K = 256;                  // entries per piece of a split line segment
  int4* sw_ent = nullptr;
  int64_t n_sw_ent = 0;
  float* psi_rot = nullptr;
  std::vector<int4> h_sw_ent;         // host copy (the weighted field is refilled by cgs_set_weights)
  // binned gridding tables
  bool use_grid_bin = false;          // binned LDS scatter (not with CGS_DETERMINISTIC)
  bool bin_allowed = false;           // build the bin tables (CGS_GRID_BIN=1)
  int bin_TB = 0, bin_cg = 0, bin_item_max = 0, bin_threads = 1024;
  int* bin_ids = nullptr;
  int4* bin_items = nullptr;
  int2* bin_range = nullptr;
  float* bin_rec = nullptr;
  float2* bin_reg = nullptr;
  int n_bin_items = 0;
  int64_t bin_nbins = 0;
  int bin_nb[3] = {1, 1, 1};

  // work buffers
  float2* grid = nullptr;             // F*ntot*C
  float2* t1 = nullptr;               // FFT intermediate
  float2* t2 = nullptr;               // second intermediate (3D)
  float2* Xc = nullptr;               // F*Ntot*C
  float2* Xc2 = nullptr;              // CGNR: f_hat
  float2* Xc3 = nullptr;              // CGNR: p_hat
  float2* ybuf[4] = {nullptr, nullptr, nullptr, nullptr};   // Mtot*C each
  float2* img[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};  // F*Ntot each
  float2* ext_a = nullptr;            // staging for host I/O
  float2* ext_b = nullptr;
  size_t ext_a_bytes = 0, ext_b_bytes = 0;
  double* partials = nullptr;         // reduction partials
  size_t n_partials = 0;
  CgState* cgst = nullptr;            // F
  int* tickets = nullptr;             // 2F last-block tickets (self-resetting)
  double* cnr = nullptr;              // CGNR per (frame, coil) scalars: 8*F*C
  int* fdone = nullptr;               // F frame-done flags
  int* fmask = nullptr;               // F scratch flags (single-frame NFFT calls)
  double* hist = nullptr;             // residual history [F][maxit+1]
  int hist_cap = 0;
  // Toeplitz embedding of E^H W E (kernel spectrum per frame)
  bool have_toep = false;
  float* That = nullptr;              // F * prod(2 N_a), conv-pass layout
  float2* tt1 = nullptr;              // (C, 2N0, N1[, N2]) per frame
  float2* tt2 = nullptr;              // (C, 2N0, 2N1, N2) per frame (3D)
  bool tt_alias = false;              // tt1/tt2 alias t1/t2 (sigma == 2)
  std::vector<float> h_k;             // trajectory copy (host)
  std::vector<float> h_w;             // weights copy (host, original order)
  CgState* h_cgst = nullptr;          // pinned mirror
  int* h_done = nullptr;              // pinned

  // off-resonance (B0) model of the NFFT operators (cgs_set_b0): a helper 3D
  // plan over (N0, N1, N3) with nodes (k_j, tau_j), per-pixel phase
  // exp(-2 pi i b0 ts) and third-axis position nu = N3 b0 / omega, per-sample
  // 1 / phihat(tau_j) in the helper's sample order
  Plan* b0q = nullptr;
  float2* b0_phase = nullptr;
  float* b0_nu = nullptr;
  float* b0_phinv = nullptr;
  int b0_N3 = 0;
  float b0_b3 = 0.f;
  double b0_ts = 0.0, b0_omega = 0.0;
  std::vector<float> h_b0, h_t;       // host copies (the helper is rebuilt by a new trajectory)

  // solve bookkeeping without allocation in the call: polling events and
  // per-frame completion times made at plan creation; the history the last
  // solve wrote (the plan's buffer, or the allocation-free solve's scratch)
  hipEvent_t poll_ev[2] = {nullptr, nullptr};
  std::vector<double> done_ms;
  double* last_hist = nullptr;
  int last_hist_cap = 0;
  bool no_prof = false;               // no timing events (allocation-free solve)
  int device = 0;

  // captured CG iterations (replayed by graph-mode solves)
  bool capturing = false;
  hipGraphExec_t graph_exec = nullptr;
  std::string graph_key;

  // profiling
  struct Stat { int64_t launches = 0; double ms = 0, bytes = 0; };
  Stat stat[CGS_K_COUNT];
  unsigned prof_mask = ~0u;           // kernel classes timed with events
  double* tail_pp = nullptr;          // fused CG tail partials (Toeplitz path)
  bool beta_folded = false;           // the captured / last CG iteration forms beta in the pad pass
  bool graph_folded = false;          // beta_folded of the iterations captured in graph_exec
  double* tail_pq = nullptr;
  // CG tail folded into the next pad pass (2D row kernels; CGS_CG_TAILFOLD=0 turns it off)
  bool tailfold = true;
  bool tailfold_used = false;         // the last solve's iterations ran folded
  bool graph_tailfold = false;        // the captured graph's iterations are folded
  double* fold_rq = nullptr;          // crop partials <r, q>, |q|^2 per (frame, row)
  double* fold_qq = nullptr;
  double* fold_rr = nullptr;          // crop partials ||r||^2 per (frame, row)
  struct Pending { hipEvent_t a, b; int kind; double bytes; };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> event_pool;
};

// ---- launchers (kernels_*.hip) ----------------------------------------------
// Kernel-timing hand-off: when set, the next FFT launch is issued with
// hipExtLaunchKernelGGL and these start/stop events (timestamps of the dispatch
// itself, no marker packets around it), and `consumed` is set.
struct LaunchEvents {
  hipEvent_t a = nullptr, b = nullptr;
  bool consumed = false;
};
LaunchEvents& launch_events();

#define CGS_LAUNCH(kernel, grid, block, shm, stream, ...)                                      \
  do {                                                                                          \
    ::cgs::LaunchEvents& le_ = ::cgs::launch_events();                                          \
    if (le_.a) {                                                                                \
      hipExtLaunchKernelGGL(kernel, grid, block, shm, stream, le_.a, le_.b, 0, __VA_ARGS__);    \
      le_.a = le_.b = nullptr;                                                                  \
      le_.consumed = true;                                                                      \
    } else {                                                                                    \
      hipLaunchKernelGGL(kernel, grid, block, shm, stream, __VA_ARGS__);                        \
    }                                                                                           \
  } while (0)

void launch_fft(const FftArgs& a, int ld_mode, int st_mode, hipStream_t s);
bool fft_reg_supported(const FftArgs& a, int ld_mode, int st_mode);
void launch_fft_reg(const FftArgs& a, int ld_mode, int st_mode, hipStream_t s);
bool toep_pad_crop_supported(int N0, int N1, int C);
bool toep_conv_supported(int N0, int N1, int C);
bool toep_half_supported(int N0, int N1, int C);
void launch_toep_half(const ToepArgs& a, int dir, hipStream_t s);
void launch_toep_pad(const ToepArgs& a, hipStream_t s);
void launch_toep_conv(const ToepArgs& a, hipStream_t s);
void launch_toep_crop(const ToepArgs& a, hipStream_t s);
void launch_interp(const SampArgs& a, int cg, hipStream_t s);
void launch_grid(const GridArgs& a, int cg, int waves, int nitems, int ncgroups, hipStream_t s);
void launch_grid_reduce(const Plan& p, const float2* partial, float2* grid, const int* fl, hipStream_t s);
void launch_grid_batched(const GridArgs& a, int cg, int waves, int nitems, int ncgroups, hipStream_t s);
void launch_interp_lds(const SampArgs& a, const int4* items, int nitems, int T, int nt0, int cg, int maxs,
                       int ncgroups, int lanes, hipStream_t s);
size_t grid_batched_lds(int d, int T, int taps, int CG, int rowstride);
void launch_grid_atomic(const GridArgs& a, const int4* vrec, int cg, int nitems, int ncgroups, hipStream_t s);
void launch_grid_reg(const GridArgs& a, const int4* vrec, int cgr, int nitems, int ncgroups, hipStream_t s);
void launch_grid_mfma(const GridArgs& a, const int4* vrec, int nitems, hipStream_t s);
void launch_grid_mfma3(const GridArgs& a, const int4* vrec, int nitems, hipStream_t s);
int sweep_groups(int C);
bool grid_sweep_supported(int taps, int C);
void launch_grid_sweep(const SweepArgs& a, hipStream_t s);
void launch_grid_sweep_reduce(const SweepArgs& a, hipStream_t s);
size_t interp_lds3_bytes(int T, int taps, int cg, int maxs);
void launch_interp_cs3(const SampArgs& a, const int4* items, int nitems, int T, int nt0, int nt1, int threads,
                       int split, hipStream_t s);
void launch_interp_lds3(const SampArgs& a, const int4* items, int nitems, int T, int nt0, int nt1, int cg, int maxs,
                        int threads, hipStream_t s);
void launch_grid_reg3(const GridArgs& a, const int4* vrec, int cgr, int nitems, int ncgroups, hipStream_t s);
size_t interp_lds_bytes(int T, int taps, int CG, int rowstride, int maxs);
void launch_cg_tail(const Plan& p, float2* x, float2* r, const float2* pv, const float2* part, int G,
                    float lambda, const double* pq_part, int npq, const double* pp_part, int npp, bool snap,
                    hipStream_t s);
const double* cg_tail_rpartials(const Plan& p);
int cg_tail_blocks(const Plan& p);
void launch_partial_combine(const Plan& p, const float2* part, int G, float2* q, const float2* pvec, float lambda,
                            double* partials, const int* fl, hipStream_t s);
void launch_coil_combine(const Plan& p, const float2* Xc, float2* x, const float2* pvec,
                         float lambda, bool use_imap, double* partials, const int* fl, hipStream_t s);
// vector / scalar kernels
void launch_sqnorm_partials(const Plan& p, const float2* v, double* partials, hipStream_t s);
void launch_cg_init(const Plan& p, const double* part_bn, const double* part_rn, int tolkind,
                    double tol, int maxit, bool warm, bool joint, hipStream_t s);
void launch_cg_alpha(const Plan& p, const double* partials, bool joint, hipStream_t s);
void launch_cg_xr(const Plan& p, float2* x, float2* r, const float2* pv, const float2* q,
                  double* partials, hipStream_t s);
void launch_cg_beta(const Plan& p, const double* partials, int tolkind, double tol, int maxit, bool joint,
                    hipStream_t s);
void launch_cg_p(const Plan& p, float2* pv, const float2* r, hipStream_t s);
void launch_coil_alpha(const Plan& p, const float2* Xc, float2* q, const float2* pv, float lambda,
                       hipStream_t s);
void launch_cg_xr_beta(const Plan& p, float2* x, float2* r, const float2* pv, const float2* q, int tolkind,
                       double tol, int maxit, hipStream_t s);
void launch_cg_xr_alpha(const Plan& p, float2* x, float2* r, const float2* pv, const float2* q, hipStream_t s);
void launch_cg_beta_after_xr(const Plan& p, int nparts, int tolkind, double tol, int maxit, hipStream_t s);
void launch_axpby_img(const Plan& p, float2* out, const float2* a, float alpha, const float2* b,
                      float beta, hipStream_t s);
// layout conversions
void launch_y_ext2int(const Plan& p, const float2* ext, float2* in, int batch, int only_frame,
                      hipStream_t s);
void launch_y_int2ext(const Plan& p, const float2* in, float2* ext, int batch, int only_frame,
                      hipStream_t s);
void launch_img_ext2coil(const Plan& p, const float2* ext, float2* X, int batch, int frame,
                         hipStream_t s);
void launch_img_coil2ext(const Plan& p, const float2* X, float2* ext, int batch, int frame,
                         hipStream_t s);
// CGNR (reference E^H) kernels
// step 1 / 2: also take the CGNR scalar step that reads this dot product (alpha / beta)
void launch_cnr_dot_samples(const Plan& p, const float2* v, double* out_slot, hipStream_t s, int step = -1);
void launch_cnr_dot_image(const Plan& p, const float2* v, bool masked, double* out_slot,
                          hipStream_t s, int step = -1);
void launch_cnr_scalars(const Plan& p, int step, double eps, hipStream_t s);
void launch_cnr_update_image(const Plan& p, float2* fhat, float2* phat, const float2* z,
                             int step, hipStream_t s);
void launch_cnr_update_samples_dot(const Plan& p, const float2* rin, float2* r, const float2* v, hipStream_t s);
void launch_copy(void* dst, const void* src, size_t bytes, hipStream_t s);
void launch_toep_layout(const float2* Khat, float* That, int64_t cells, int nl, int F, float scale,
                        hipStream_t s);
void launch_fill_int(int* dst, int value, int n, hipStream_t s);
// off-resonance model (kernels_b0.hip)
void launch_b0_spread(const float2* x, const float2* phase, const float* nu, float2* X3, int64_t Ntot, int C, int N3,
                      int m, float b3, int mask_nh, int N0, int N1, hipStream_t s);
void launch_b0_gather(const float2* X3, const float2* phase, const float* nu, float2* x, int64_t Ntot, int C, int N3,
                      int m, float b3, hipStream_t s);
void launch_b0_scale(const float2* y, const float* scale, float2* out, int64_t M, int C, hipStream_t s);
void launch_zero(void* dst, size_t bytes, hipStream_t s);   // bytes a multiple of 4
void launch_grid_bin(const BinArgs& a, int threads, hipStream_t s);
void launch_grid_bin_sum(const BinArgs& a, hipStream_t s);
size_t grid_bin_lds(int d, int taps, int CG, int E, int P1);
int bin_row_stride(int E);
void launch_cg_fold_final(const Plan& p, float2* x, float2* r, const float2* pv, const float2* qv, const double* fpq,
                          const double* frq, const double* fqq, const double* frr, int nrpart, int tolkind,
                          double tol, int maxit,
                          hipStream_t s);
void launch_maps_load(const float2* S, float2* Sd, float* I, int64_t Ntot, int64_t nm, int C, hipStream_t s);

// read-only tables that a kernel reads at wave-uniform addresses: through the
// constant address space the compiler issues scalar loads (s_load) for them
#if defined(__HIP_DEVICE_COMPILE__)
template <class T>
using const_ptr = const __attribute__((address_space(4))) T*;
#else   // the host pass only needs the kernels' signatures
template <class T>
using const_ptr = const T*;
#endif
template <class T>
__device__ __forceinline__ const_ptr<T> as_const(const T* p) {
  return (const_ptr<T>)(p);
}

// errors
void set_error(const std::string& msg);

}  // namespace cgs
