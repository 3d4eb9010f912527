// plan.hip -- host side of libcgsense_hip: plan construction, trajectory
// presort, operator composition (E, E^H, CGNR, CG) and the C ABI.
//
// Reference counterparts (codeare source tree):
//   plan        NFFT<T>::NFFT(Params)       src/matrix/ft/NFFT.hpp:66-191
//               NCSENSE<T>::NCSENSE(Params) src/matrix/ft/NCSENSE.hpp:69-184
//   trajectory  NFFT::KSpace / Weights      NFFT.hpp:258-296, NCSENSE.hpp:204-248
//   E           NCSENSE::Trafo              NCSENSE.hpp:310-348
//   E^H (ref)   NCSENSE::operator/          NCSENSE.hpp:251-301 -> NFFT::Adjoint (ITrafo)
//   solve       NCSENSE::Adjoint -> CGLS    NCSENSE.hpp:408-414, CGLS.hpp:30-59
// One plan replaces the C per-coil NFFT3 plans of NCSENSE: the trajectory and
// window are stored once and every kernel is batched over coils and frames.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <cstdlib>
#include <chrono>
#include <vector>

#include "cgs_internal.hpp"

struct cgs_plan_s {
  cgs::Plan p;
};

namespace cgs {

static thread_local std::string g_err;
void set_error(const std::string& m) { g_err = m; }

LaunchEvents& launch_events() {
  static thread_local LaunchEvents le;
  return le;
}

#define HIPCHK(x)                                                                    \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess) {                                                          \
      set_error(std::string(#x) + " failed: " + hipGetErrorString(e_));              \
      return CGS_ERR_HIP;                                                            \
    }                                                                                \
  } while (0)

#define CHK(x)                     \
  do {                             \
    int r_ = (x);                  \
    if (r_ != CGS_OK) return r_;   \
  } while (0)

// ---- math for tables (FP64) --------------------------------------------------
static double bessel_i0(double x) {
  // power series sum_k (x^2/4)^k / (k!)^2
  double q = 0.25 * x * x, term = 1.0, sum = 1.0;
  for (int k = 1; k < 500; ++k) {
    term *= q / ((double)k * k);
    sum += term;
    if (term < 1e-17 * sum) break;
  }
  return sum;
}

// Kaiser-Bessel window of NFFT3 in grid-cell units (NFFTTraits.hpp:665-673),
// b = pi (2 - 1/sigma); continuous through |t| = m where it equals b/pi.
static double kb_phi(double t, int m, double b) {
  const double s2 = (double)m * m - t * t;
  if (s2 > 0) { const double s = std::sqrt(s2); return std::sinh(b * s) / (M_PI * s); }
  if (s2 < 0) { const double s = std::sqrt(-s2); return std::sin(b * s) / (M_PI * s); }
  return b / M_PI;
}

static bool factorize(int n, int* radix, int* nst);

static int make_tab(FftTab& t, int n);

static bool factorize(int n, int* radix, int* nst) {
  int c2 = 0, r = n;
  while (r % 2 == 0) { r /= 2; ++c2; }
  int c3 = 0;
  while (r % 3 == 0) { r /= 3; ++c3; }
  if (r != 1) return false;
  int k = 0;
  while (c2 >= 3) { radix[k++] = 8; c2 -= 3; }
  if (c2 == 2) radix[k++] = 4;
  if (c2 == 1) radix[k++] = 2;
  while (c3-- > 0) radix[k++] = 3;
  if (k > 8) return false;
  *nst = k;
  return true;
}

template <typename T>
static int dalloc(T** ptr, size_t count) {
  *ptr = nullptr;
  if (count == 0) return CGS_OK;
  hipError_t e = hipMalloc((void**)ptr, sizeof(T) * count);
  if (e != hipSuccess) {
    set_error(std::string("hipMalloc of ") + std::to_string(sizeof(T) * count) + " bytes failed: " +
              hipGetErrorString(e));
    *ptr = nullptr;
    return CGS_ERR_ALLOC;
  }
  return CGS_OK;
}
template <typename T>
static void dfree(T*& ptr) {
  if (ptr) (void)hipFree((void*)ptr);
  ptr = nullptr;
}

static int make_tab(FftTab& t, int n) {
  t.n = n;
  if (!factorize(n, t.radix, &t.nst)) {
    set_error("FFT length " + std::to_string(n) + " is not 2^a 3^b");
    return CGS_ERR_UNSUPPORTED;
  }
  std::vector<float2> tw(n);
  for (int q = 0; q < n; ++q) {
    const double ang = -2.0 * M_PI * (double)q / (double)n;
    tw[q] = make_float2((float)std::cos(ang), (float)std::sin(ang));
  }
  CHK(dalloc(&t.tw, tw.size()));
  HIPCHK(hipMemcpy(t.tw, tw.data(), sizeof(float2) * tw.size(), hipMemcpyHostToDevice));
  return CGS_OK;
}

// ---- profiling -----------------------------------------------------------------
struct Prof {
  Plan& p;
  int kind;
  double bytes;
  hipEvent_t a = nullptr, b = nullptr;
  bool timed = false, ext = false;
  // no timing events while a HIP graph is being captured; classes outside
  // prof_mask are counted (launches, bytes) but not timed.  FFT passes (one
  // launch per scope) hand the events to the launch itself
  // (hipExtLaunchKernelGGL), which timestamps the dispatch; other scopes
  // bracket their launches with event records.
  static bool fft_kind(int k) {
    return k == CGS_K_FFT_FWD || k == CGS_K_FFT_ADJ || k == CGS_K_TOEP_PAD || k == CGS_K_TOEP_CONV ||
           k == CGS_K_TOEP_CROP;
  }
  Prof(Plan& p_, int kind_, double bytes_) : p(p_), kind(kind_), bytes(bytes_) {
    if (!(p.flags & CGS_PROFILE) || p.capturing || p.no_prof) return;
    if (!((p.prof_mask >> kind) & 1u)) {
      p.stat[kind].launches += 1;
      p.stat[kind].bytes += bytes;
      return;
    }
    timed = true;
    a = take();
    b = take();
    if (fft_kind(kind)) {
      ext = true;
      LaunchEvents& le = launch_events();
      le.a = a;
      le.b = b;
      le.consumed = false;
    } else {
      (void)hipEventRecord(a, p.stream);
    }
  }
  ~Prof() {
    if (!timed) return;
    if (ext) {
      LaunchEvents& le = launch_events();
      if (!le.consumed) {   // no FFT launch happened in the scope
        (void)hipEventRecord(a, p.stream);
        (void)hipEventRecord(b, p.stream);
      }
      le = LaunchEvents{};
    } else {
      (void)hipEventRecord(b, p.stream);
    }
    p.pending.push_back({a, b, kind, bytes});
  }
  hipEvent_t take() {
    if (!p.event_pool.empty()) {
      hipEvent_t e = p.event_pool.back();
      p.event_pool.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
  }
};

static int drain_profile(Plan& p) {
  if (p.pending.empty()) return CGS_OK;
  HIPCHK(hipStreamSynchronize(p.stream));
  for (auto& q : p.pending) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, q.a, q.b);
    p.stat[q.kind].launches += 1;
    p.stat[q.kind].ms += ms;
    p.stat[q.kind].bytes += q.bytes;
    p.event_pool.push_back(q.a);
    p.event_pool.push_back(q.b);
  }
  p.pending.clear();
  return CGS_OK;
}

// ---- FFT passes ------------------------------------------------------------------
// lines staged per FFT block; CGS_FFT_LB caps it (read once per process)
static int fft_lb_cap() {
  static const int cap = [] {
    const char* e = std::getenv("CGS_FFT_LB");
    const int v = e ? std::atoi(e) : 64;
    return (v >= 1 && v <= 64) ? v : 64;
  }();
  return cap;
}

static int pick_lb(int64_t inner, int n) {
  int lb = 1;
  const int cap = fft_lb_cap();
  while (lb * 2 <= kFftMaxElem / n && lb * 2 <= inner && lb * 2 <= cap) lb *= 2;
  return lb;
}

static FftArgs fft_base(const Plan& p, const FftTab& tab, int sign) {
  FftArgs a{};
  a.n = tab.n;
  a.sign = sign;
  a.pad_in = sign < 0;    // forward passes read image-index order
  a.crop_out = sign > 0;  // adjoint passes write image-index order
  a.tw = tab.tw;
  a.nst = tab.nst;
  for (int k = 0; k < 8; ++k) a.radix[k] = tab.radix[k];
  a.dinv0 = p.dinv[0];
  a.dinv1 = p.dinv[1];
  a.dinv2 = p.d == 3 ? p.dinv[2] : p.dinv[1];
  a.C = p.C;
  a.d = p.d;
  a.N0 = p.N[0];
  a.N1 = p.N[1];
  a.Ntot = p.Ntot;
  a.S = p.S;
  a.Sfs = p.Sfs;
  a.mask_nh = (int)(0.5 * p.N[0]);
  return a;
}

static bool fft_reg_enabled() {
  static const bool on = [] {
    const char* e = std::getenv("CGS_FFT_REG");
    return !(e && std::atoi(e) == 0);
  }();
  return on;
}

static void fft_finish(FftArgs& a) {
  a.use_reg = fft_reg_enabled() && fft_reg_supported(a, 0, 0);
  a.lb = a.use_reg ? 16 : pick_lb(a.inner, a.n);
  a.lb_log2 = 0;
  while ((1 << a.lb_log2) < a.lb) ++a.lb_log2;
  a.ldn = a.n + (a.n >> 4) + 1;   // room for the 1-in-16 pad slots of the line layout
  uint64_t Ns = 1;
  for (int s = 0; s < a.nst; ++s) {
    a.ns_magic[s] = (Ns == 1) ? 0u : (uint32_t)(((1ull << 32) + Ns - 1) / Ns);
    Ns *= (uint64_t)a.radix[s];
  }
}

// fused CG direction update carried into the first FFT pass (p_new = r + beta p_old)
struct PUpd {
  const float2* r = nullptr;
  const float2* pold = nullptr;
  float2* pnew = nullptr;
};

static void apply_pupd(const Plan& p, FftArgs& a, const PUpd* pu) {
  if (!pu) return;
  a.r = pu->r;
  a.pold = pu->pold;
  a.pnew = pu->pnew;
  a.cgst = p.cgst;
}

// image (mode 1: x .* S_c; 2: X; 3: X .* w_hat) -> grid, forward transform
static int fft_forward(Plan& p, int mode, const float2* src, const int* fl, const PUpd* pu = nullptr) {
  const int d = p.d, C = p.C, F = p.F;
  const int64_t N0 = p.N[0], N1 = p.N[1], N2 = p.N[2], n0 = p.n[0], n1 = p.n[1], n2 = p.n[2];
  double bytes = (mode == 1 ? (double)(1 + C) * p.Ntot * 8 : 2.0 * C * p.Ntot * 8) * F;
  // pass 1, axis 0
  {
    FftArgs a = fft_base(p, p.fa[0], -1);
    a.in = mode == 1 ? nullptr : src;
    a.x = mode == 1 ? src : nullptr;
    if (mode == 1) apply_pupd(p, a, pu);
    a.out = (d == 2) ? p.t1 : p.t1;
    a.nin = (int)N0; a.nout = (int)n0;
    a.inner = C;
    a.outer_per_frame = (d == 2) ? N1 : N1 * N2;
    a.outers = a.outer_per_frame * F;
    a.fdone = fl;
    fft_finish(a);
    const double b = bytes + (double)C * n0 * a.outer_per_frame * F * 8;
    Prof pr(p, CGS_K_FFT_FWD, b);
    launch_fft(a, mode, 0, p.stream);
  }
  // pass 2, axis 1
  {
    FftArgs a = fft_base(p, p.fa[1], -1);
    a.in = p.t1;
    a.out = (d == 2) ? p.grid : p.t2;
    a.nin = (int)N1; a.nout = (int)n1;
    a.inner = C * n0;
    a.outer_per_frame = (d == 2) ? 1 : N2;
    a.outers = a.outer_per_frame * F;
    a.fdone = fl;
    fft_finish(a);
    const double b = (double)C * n0 * (N1 + n1) * a.outer_per_frame * F * 8;
    Prof pr(p, CGS_K_FFT_FWD, b);
    launch_fft(a, 0, 0, p.stream);
  }
  if (d == 3) {
    FftArgs a = fft_base(p, p.fa[2], -1);
    a.in = p.t2;
    a.out = p.grid;
    a.nin = (int)N2; a.nout = (int)n2;
    a.inner = C * n0 * n1;
    a.outer_per_frame = 1;
    a.outers = F;
    a.fdone = fl;
    fft_finish(a);
    const double b = (double)C * n0 * n1 * (N2 + n2) * F * 8;
    Prof pr(p, CGS_K_FFT_FWD, b);
    launch_fft(a, 0, 0, p.stream);
  }
  return CGS_OK;
}

// grid -> per-coil images X (cropped, deapodised), adjoint transform
static int fft_adjoint(Plan& p, float2* X, const int* fl) {
  const int d = p.d, C = p.C, F = p.F;
  const int64_t N0 = p.N[0], N1 = p.N[1], N2 = p.N[2], n0 = p.n[0], n1 = p.n[1], n2 = p.n[2];
  if (d == 2) {
    {
      FftArgs a = fft_base(p, p.fa[1], +1);
      a.in = p.grid; a.out = p.t1;
      a.nin = (int)n1; a.nout = (int)N1;
      a.inner = C * n0; a.outer_per_frame = 1; a.outers = F; a.fdone = fl;
      fft_finish(a);
      Prof pr(p, CGS_K_FFT_ADJ, (double)C * n0 * (n1 + N1) * F * 8);
      launch_fft(a, 0, 0, p.stream);
    }
    {
      FftArgs a = fft_base(p, p.fa[0], +1);
      a.in = p.t1; a.out = X;
      a.nin = (int)n0; a.nout = (int)N0;
      a.inner = C; a.outer_per_frame = N1; a.outers = N1 * F; a.fdone = fl;
      fft_finish(a);
      Prof pr(p, CGS_K_FFT_ADJ, (double)C * (n0 + N0) * N1 * F * 8);
      launch_fft(a, 0, 1, p.stream);
    }
  } else {
    {
      FftArgs a = fft_base(p, p.fa[2], +1);
      a.in = p.grid; a.out = p.t2;
      a.nin = (int)n2; a.nout = (int)N2;
      a.inner = C * n0 * n1; a.outer_per_frame = 1; a.outers = F; a.fdone = fl;
      fft_finish(a);
      Prof pr(p, CGS_K_FFT_ADJ, (double)C * n0 * n1 * (n2 + N2) * F * 8);
      launch_fft(a, 0, 0, p.stream);
    }
    {
      FftArgs a = fft_base(p, p.fa[1], +1);
      a.in = p.t2; a.out = p.t1;
      a.nin = (int)n1; a.nout = (int)N1;
      a.inner = C * n0; a.outer_per_frame = N2; a.outers = N2 * F; a.fdone = fl;
      fft_finish(a);
      Prof pr(p, CGS_K_FFT_ADJ, (double)C * n0 * (n1 + N1) * N2 * F * 8);
      launch_fft(a, 0, 0, p.stream);
    }
    {
      FftArgs a = fft_base(p, p.fa[0], +1);
      a.in = p.t1; a.out = X;
      a.nin = (int)n0; a.nout = (int)N0;
      a.inner = C; a.outer_per_frame = N1 * N2; a.outers = N1 * N2 * F; a.fdone = fl;
      fft_finish(a);
      Prof pr(p, CGS_K_FFT_ADJ, (double)C * (n0 + N0) * N1 * N2 * F * 8);
      launch_fft(a, 0, 1, p.stream);
    }
  }
  return CGS_OK;
}

// ---- window gather / scatter --------------------------------------------------------
constexpr int kInterp3Batch = 64;   // samples per staged batch of the 3D LDS gather
static int interp(Plan& p, float2* y, const int* fl) {
  SampArgs a{};
  a.base = p.base; a.psi = p.psi; a.w = nullptr;
  a.d = p.d; a.taps = p.taps; a.C = p.C;
  a.n0 = p.n[0]; a.n1 = p.n[1]; a.n2 = p.n[2];
  a.ntot = p.ntot;
  a.grid = p.grid; a.y = y;
  a.items = p.interp_items; a.nitems = p.n_interp_items;
  a.fdone = fl;
  const double bytes = (double)p.Mtot * p.d * 4 + (double)p.C * p.ntot * p.F * 8 + (double)p.C * p.Mtot * 8;
  Prof pr(p, CGS_K_INTERP, bytes);
  if (p.d == 2 && p.use_interp_lds) {
    const int ng = (p.C + p.interp_cg - 1) / p.interp_cg;
    launch_interp_lds(a, p.tile_items, p.n_tile_items, p.T, p.nt[0], p.interp_cg, p.tile_item_max, ng,
                      p.interp_lanes, p.stream);
  } else if (p.d == 3 && p.use_interp_lds && p.interp_lanes >= 1 && p.taps == 8 && p.T == 8) {
    // (8 + 7)^3 cells x 2 coils staged once per item (54 KB), lanes = (sample, tap plane[, tap parity])
    launch_interp_cs3(a, p.tile_items, p.n_tile_items, p.T, p.nt[0], p.nt[1], p.interp3_threads,
                      p.interp_lanes == 2 ? 1 : 0, p.stream);
  } else if (p.d == 3 && p.use_interp_lds && p.T == 8 && p.taps <= 8 &&
             interp_lds3_bytes(p.T, p.taps, 2, kInterp3Batch) <= 65536) {
    // (8 + taps - 1)^3 cells x 2 coils staged once per item (54 KB at taps
    // 8), its samples taken kInterp3Batch at a time
    launch_interp_lds3(a, p.tile_items, p.n_tile_items, p.T, p.nt[0], p.nt[1], 2, kInterp3Batch, p.interp3_threads,
                       p.stream);
  } else {
    launch_interp(a, p.interp_cg, p.stream);
  }
  return CGS_OK;
}

static void bin_args(const Plan& p, const float2* y, bool weighted, const int* fl, BinArgs& a);
static int build_bins(Plan& p, const std::vector<int>& hbase, const std::vector<float>& hpsi,
                      const std::vector<int>& hframe);

static int gridding(Plan& p, const float2* y, bool weighted, const int* fl) {
  GridArgs a{};
  a.base = p.base; a.psi = p.psi; a.w = weighted ? p.w : nullptr;
  a.y = y; a.grid = p.grid; a.partial = p.partial;
  a.items = p.grid_items; a.visits = p.visits; a.fdone = fl;
  a.d = p.d; a.taps = p.taps; a.C = p.C; a.T = p.T;
  const bool batched = p.use_grid_batched;
  a.CGW = batched ? p.grid_cg : p.grid_cg * p.grid_w;
  a.rowstride = p.T * a.CGW + ((a.CGW == 8) ? 8 : 0);
  a.n0 = p.n[0]; a.n1 = p.n[1]; a.n2 = p.n[2];
  a.nt0 = p.nt[0]; a.nt1 = p.nt[1]; a.nt2 = p.nt[2];
  a.ntot = p.ntot;
  a.tilecells = (int64_t)p.T * p.T * (p.d == 3 ? p.T : 1);
  const int ngroups = (p.C + a.CGW - 1) / a.CGW;
  const double bytes = (double)p.Mtot * p.d * 4 + (double)p.Mtot * 4 + (double)p.C * p.Mtot * 8 +
                       (double)p.C * p.ntot * p.F * 8;
  Prof pr(p, CGS_K_GRID, bytes);
  if (p.use_grid_bin && p.bin_ids) {
    BinArgs ba;
    bin_args(p, y, weighted, fl, ba);
    launch_grid_bin(ba, p.bin_threads, p.stream);
    launch_grid_bin_sum(ba, p.stream);
    return CGS_OK;
  }
  if (p.d == 2 && p.use_grid_sweep && p.sw_items) {
    SweepArgs sa{};
    sa.items = p.sw_items; sa.ent = p.sw_ent; sa.psi_rot = p.psi_rot;
    sa.y = y; sa.grid = p.grid; sa.part = p.sw_part; sa.fdone = fl;
    sa.ritems = p.sw_ritems; sa.rpieces = p.sw_rpieces; sa.nritems = p.n_sw_ritems;
    sa.nitems = p.n_sw_items; sa.taps = p.taps; sa.C = p.C;
    sa.W = std::max(1, p.sweep_W / p.taps) * p.taps;
    sa.nseg = p.sw_nseg; sa.n0 = p.n[0]; sa.weighted = weighted ? 1 : 0;
    sa.ntot = p.ntot;
    launch_grid_sweep(sa, p.stream);
    launch_grid_sweep_reduce(sa, p.stream);
    return CGS_OK;
  }
  if (p.d == 2 && p.T == 8 && p.use_grid_mfma) {
    launch_grid_mfma(a, p.vrec, p.n_grid_items, p.stream);
  } else if (p.d == 2 && p.T == 8 && p.use_grid_reg) {
    const int cgr = p.grid_cgr;
    launch_grid_reg(a, p.vrec, cgr, p.n_grid_items, (p.C + cgr - 1) / cgr, p.stream);
  } else if (p.d == 3 && p.T == 8 && p.use_grid_mfma) {
    launch_grid_mfma3(a, p.vrec, p.n_grid_items, p.stream);
  } else if (p.d == 3 && p.T == 8 && p.use_grid_reg) {
    const int cgr = std::min(p.grid_cgr, 16);
    launch_grid_reg3(a, p.vrec, cgr, p.n_grid_items, (p.C + cgr - 1) / cgr, p.stream);
  } else if (batched && p.use_grid_atomic)
    launch_grid_atomic(a, p.vrec, p.grid_cg, p.n_grid_items, ngroups, p.stream);
  else if (batched)
    launch_grid_batched(a, p.grid_cg, p.grid_w, p.n_grid_items, ngroups, p.stream);
  else
    launch_grid(a, p.grid_cg, p.grid_w, p.n_grid_items, ngroups, p.stream);
  launch_grid_reduce(p, p.partial, p.grid, fl, p.stream);
  return CGS_OK;
}

static double coil_bytes(const Plan& p, bool lam) {
  return ((double)2 * p.C * p.Ntot * 8 + (double)p.Ntot * 8 * (lam ? 2 : 1)) * p.F;
}

// ---- NCSENSE operators ----------------------------------------------------------------
// E: y_int = B F D (S_c .* x) for all coils and frames
static int op_E(Plan& p, const float2* x, float2* y, const int* fl, const PUpd* pu = nullptr) {
  CHK(fft_forward(p, 1, x, fl, pu));
  CHK(interp(p, y, fl));
  return CGS_OK;
}

// true E^H: x = sum_c conj(S_c) D F^H B^T (w .* y_c) [+ lambda pv, partial <pv, x>]
static int op_EH_true(Plan& p, const float2* y, float2* x, const float2* pv, float lambda,
                      double* partials, const int* fl) {
  CHK(gridding(p, y, true, fl));
  CHK(fft_adjoint(p, p.Xc, fl));
  Prof pr(p, CGS_K_COIL, coil_bytes(p, pv != nullptr));
  launch_coil_combine(p, p.Xc, x, pv, lambda, false, partials, fl, p.stream);
  return CGS_OK;
}

// The CGNR's operator pair: A^H W (samples -> p.Xc) and A w_hat (image -> samples).
// Default: the plan's NFFT (gridding + adjoint FFT passes, masked forward FFT
// passes + gather); the off-resonance model supplies its own.
struct CnrOps {
  std::function<int(const float2* r)> adj_w;
  std::function<int(const float2* x, float2* v)> fwd_masked;
};

// NFFT3 weighted damped CGNR (NFFTTraits.hpp:122-138) for every (frame, coil):
// f_hat = ITrafo(y), written to p.Xc2.  Frames flagged in fl are skipped.
static int cgnr(Plan& p, const float2* y, int maxit, double eps, const int* fl, const CnrOps* ops = nullptr) {
  const size_t Fb = sizeof(int) * p.F;
  if (fl && fl != p.fmask) HIPCHK(hipMemcpyAsync(p.fmask, fl, Fb, hipMemcpyDeviceToDevice, p.stream));
  if (!fl) launch_fill_int(p.fmask, 0, p.F, p.stream);
  const int* m = p.fmask;
  const size_t ybytes = sizeof(float2) * (size_t)p.Mtot * p.C;
  const size_t xbytes = sizeof(float2) * (size_t)p.F * p.Ntot * p.C;
  const int FC = p.F * p.C;
  float2* r = p.ybuf[2];
  float2* v = p.ybuf[3];
  {
    // r = y - A 0 = y: the first step reads y and writes r, so y is not copied
    Prof pr(p, CGS_K_CG, 1.0 * ybytes);
    launch_cnr_scalars(p, 3, eps, p.stream);
    launch_cnr_dot_samples(p, y, p.cnr + 0 * FC, p.stream);                    // dot_r
  }
  auto adj_w = [&](const float2* rr) -> int {                                   // z = A^H W r
    if (ops) return ops->adj_w(rr);
    CHK(gridding(p, rr, true, m));
    return fft_adjoint(p, p.Xc, m);
  };
  auto fwd_masked = [&](const float2* xx, float2* vv) -> int {                 // v = A(w_hat p)
    if (ops) return ops->fwd_masked(xx, vv);
    CHK(fft_forward(p, 3, xx, m));
    return interp(p, vv, m);
  };
  CHK(adj_w(y));
  {
    Prof pr(p, CGS_K_CG, 3.0 * xbytes);
    launch_cnr_dot_image(p, p.Xc, true, p.cnr + 1 * FC, p.stream);             // dot_z
    HIPCHK(hipMemcpyAsync(p.Xc3, p.Xc, xbytes, hipMemcpyDeviceToDevice, p.stream));   // p = z
    launch_zero(p.Xc2, xbytes, p.stream);                                       // f_hat = 0
  }
  for (int l = 0; l < maxit; ++l) {
    launch_cnr_scalars(p, 0, eps, p.stream);                                   // dot_r < eps -> stop
    CHK(fwd_masked(p.Xc3, v));                                                  // v = A(w_hat p)
    {
      Prof pr(p, CGS_K_CG, 4.0 * xbytes + 4.0 * ybytes);
      launch_cnr_dot_samples(p, v, p.cnr + 2 * FC, p.stream, 1);               // dot_v; alpha
      launch_cnr_update_image(p, p.Xc2, p.Xc3, nullptr, 0, p.stream);          // f_hat += a w_hat p
      launch_cnr_update_samples_dot(p, l == 0 ? y : r, r, v, p.stream);        // r -= a v; dot_r
    }
    CHK(adj_w(r));                                                              // z = A^H W r
    {
      Prof pr(p, CGS_K_CG, 4.0 * xbytes);
      launch_cnr_dot_image(p, p.Xc, true, p.cnr + 6 * FC, p.stream, 2);        // dot_z (new); beta
      launch_cnr_update_image(p, nullptr, p.Xc3, p.Xc, 1, p.stream);           // p = b p + z
    }
  }
  return CGS_OK;
}

// reference E^H (NCSENSE::operator/): x = I .* sum_c conj(S_c) .* ITrafo_c(y_c)
static int op_EH_ref(Plan& p, const float2* y, float2* x, int ftiter, double fteps, const float2* pv,
                     float lambda, double* partials, const int* fl) {
  CHK(cgnr(p, y, ftiter, fteps, fl));
  Prof pr(p, CGS_K_COIL, coil_bytes(p, pv != nullptr));
  launch_coil_combine(p, p.Xc2, x, pv, lambda, true, partials, fl, p.stream);
  return CGS_OK;
}

static int op_EH(Plan& p, int mode, const float2* y, float2* x, int ftiter, double fteps, const float2* pv,
                 float lambda, double* partials, const int* fl) {
  if (mode == CGS_EH_REF) return op_EH_ref(p, y, x, ftiter, fteps, pv, lambda, partials, fl);
  return op_EH_true(p, y, x, pv, lambda, partials, fl);
}

static int create_plan(Plan& p, int d, const int* N, int64_t Mtot, double sigma, int m, int C, int F,
                       const int64_t* foff, unsigned flags);
static void destroy_plan(Plan& p);
static void toep_drop(Plan& p);
static int autotune(Plan& p, int reps, bool use_cache, cgs_tune_report* rep);

// ---- trajectory presort -----------------------------------------------------------------
static void free_sweep(Plan& p) {
  dfree(p.sw_items); dfree(p.sw_ent); dfree(p.psi_rot);
  dfree(p.sw_ritems); dfree(p.sw_rpieces); dfree(p.sw_part);
  p.n_sw_items = 0; p.n_sw_ent = 0; p.sw_nseg = 0; p.n_sw_ritems = 0;
  p.h_sw_ent.clear();
  p.h_sw_ent.shrink_to_fit();
}

static void free_bins(Plan& p) {
  dfree(p.bin_ids); dfree(p.bin_items); dfree(p.bin_range); dfree(p.bin_rec); dfree(p.bin_reg);
  p.n_bin_items = 0; p.bin_nbins = 0; p.bin_TB = 0;
}

// frees the tile-sorted sample tables (and, with sweep, the line-sweep and bin tables)
static void free_samples(Plan& p, bool sweep = true) {
  if (sweep) { free_sweep(p); free_bins(p); }
  dfree(p.base); dfree(p.psi); dfree(p.perm); dfree(p.iperm); dfree(p.sframe);
  dfree(p.interp_items); dfree(p.tile_items); dfree(p.grid_items);
  p.n_tile_items = 0; dfree(p.reduce_items); dfree(p.visits); dfree(p.vrec); dfree(p.partial);
  p.n_interp_items = p.n_grid_items = p.n_reduce_items = 0;
  p.n_visits = 0;
  p.n_slots = 0;
}

// Line-sweep scatter tables (2D, kernels_sweep.hip).  Work items are the
// (frame, line l1, segment [S, S + W)) triples, all of them (an item without
// entries writes its cells' zeros).  An entry is a sample whose footprint covers
// line l1 and meets the segment, with its first footprint cell r0 relative to
// S in [-(taps - 1), W): for each segment the footprint (mod n0) meets, the
// aliases b0 - S and b0 - S -/+ n0 that fall in that range (two of them only
// when the footprint wraps around a grid shorter than W + taps).  Entries are
// stable-sorted by r0 (sample order within one r0); items heaviest first.
static int build_sweep(Plan& p, const std::vector<int>& hbase, const std::vector<float>& hpsi,
                       const std::vector<int>& hframe) {
  const int taps = p.taps;
  if (p.d != 2 || !p.sweep_allowed || p.n[0] % taps != 0 || !grid_sweep_supported(taps, p.C)) return CGS_OK;
  const int n0 = p.n[0], n1 = p.n[1];
  const int W = std::max(1, p.sweep_W / taps) * taps;
  const int nseg = (n0 + W - 1) / W;
  const int64_t M = p.Mtot;
  const int64_t nitems = (int64_t)p.F * n1 * nseg;
  if (nitems > INT32_MAX / 2 || M * taps * 4 > INT32_MAX) {
    set_error("line-sweep tables exceed 32-bit indices");
    return CGS_ERR_UNSUPPORTED;
  }
  // segments and start cells of one sample's axis-0 footprint
  auto for_segments = [&](int b0, auto&& fn) {
    int sg[2], ns = 0;
    const int ends[2] = {b0, (b0 + taps - 1) % n0};
    for (int z = 0; z < 2; ++z) {
      const int v = ends[z] / W;
      if (ns == 0 || sg[0] != v) sg[ns++] = v;
    }
    for (int z = 0; z < ns; ++z) {
      const int S = sg[z] * W, E = std::min(W, n0 - S);
      const int r = b0 - S;
      const int cand[3] = {r, r - n0, r + n0};
      for (int q = 0; q < 3; ++q)
        if (cand[q] >= -(taps - 1) && cand[q] < E) fn(sg[z], cand[q]);
    }
  };
  std::vector<int64_t> cnt((size_t)nitems + 1, 0);
  for (int64_t s = 0; s < M; ++s) {
    const int b0 = hbase[s * 2], b1 = hbase[s * 2 + 1];
    const int64_t fb = (int64_t)hframe[s] * n1;
    for (int t1 = 0; t1 < taps; ++t1) {
      const int l1 = (b1 + t1) % n1;
      for_segments(b0, [&](int sg, int) { cnt[(size_t)((fb + l1) * nseg + sg) + 1]++; });
    }
  }
  for (int64_t q = 0; q < nitems; ++q) cnt[q + 1] += cnt[q];
  const int64_t nent = cnt[nitems];
  if (nent > INT32_MAX) {
    set_error("too many line-sweep entries for 32-bit indices");
    return CGS_ERR_UNSUPPORTED;
  }
  // entries grouped by item in sample order, then stable-sorted by r0 per item
  std::vector<int4> ent((size_t)nent);
  {
    std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
    for (int64_t s = 0; s < M; ++s) {
      const int b0 = hbase[s * 2], b1 = hbase[s * 2 + 1];
      const int64_t fb = (int64_t)hframe[s] * n1;
      for (int t1 = 0; t1 < taps; ++t1) {
        const int l1 = (b1 + t1) % n1;
        const float w1 = hpsi[(s * 2 + 1) * taps + t1];
        for_segments(b0, [&](int sg, int r0) {
          const float wf = w1;
          int4 e;
          e.x = (int)s;
          e.y = r0;
          std::memcpy(&e.z, &wf, 4);
          std::memcpy(&e.w, &wf, 4);
          ent[(size_t)pos[(size_t)((fb + l1) * nseg + sg)]++] = e;
        });
      }
    }
  }
  for (int64_t q = 0; q < nitems; ++q)
    std::stable_sort(ent.begin() + cnt[q], ent.begin() + cnt[q + 1],
                     [](const int4& a, const int4& b) { return a.y < b.y; });
  // Work items: a line segment with at most K entries is one item writing its
  // cells to the grid; a longer one is split into pieces of about K entries
  // (contiguous in r0 order) that write the cells they touch to partial
  // ranges, summed in piece order by the reduce pass.  Items go to the eight
  // XCDs (a block's XCD is its index mod 8) by bands of 8 lines, so that the
  // lines sharing samples share an L2; heaviest first within an XCD.
  const int K = std::max(16, p.sweep_K);
  constexpr int kXcd = 8, kWavesPerBlock = 4;
  struct Piece { int4 a, b; int64_t n; };
  std::vector<std::vector<Piece>> per_xcd(kXcd);
  std::vector<int4> ritems, rpieces;
  int64_t part_floats = 0;
  const int ncomp = 2 * p.C;
  for (int64_t q = 0; q < nitems; ++q) {
    const int f = (int)(q / ((int64_t)n1 * nseg));
    const int line = (int)(q % ((int64_t)n1 * nseg));
    const int l1 = line / nseg;
    const int E = std::min(W, n0 - (line % nseg) * W);
    const int64_t e0 = cnt[q], e1 = cnt[q + 1], ne = e1 - e0;
    auto& dst = per_xcd[(size_t)((l1 / 8) % kXcd)];
    if (ne <= K) {
      dst.push_back(Piece{make_int4((int)e0, (int)e1, f, line), make_int4(-1, 0, E, 0), ne});
      continue;
    }
    const int npc = (int)((ne + K - 1) / K);
    ritems.push_back(make_int4((int)rpieces.size(), (int)rpieces.size() + npc, f, line));
    for (int j = 0; j < npc; ++j) {
      const int64_t a = e0 + ne * j / npc, b = e0 + ne * (j + 1) / npc;
      const int ca = std::max(0, ent[(size_t)a].y), cb = std::min(E, ent[(size_t)(b - 1)].y + taps);
      const int4 pc = make_int4((int)part_floats, ca, cb, 0);
      rpieces.push_back(pc);
      dst.push_back(Piece{make_int4((int)a, (int)b, f, line), pc, b - a});
      part_floats += (int64_t)(cb - ca) * ncomp;
    }
  }
  if (part_floats > INT32_MAX) {
    set_error("line-sweep partial ranges exceed 32-bit offsets");
    return CGS_ERR_UNSUPPORTED;
  }
  size_t per_max = 0;
  for (auto& v : per_xcd) {
    std::stable_sort(v.begin(), v.end(), [](const Piece& x, const Piece& y) { return x.n > y.n; });
    per_max = std::max(per_max, v.size());
  }
  // item i runs in block i / 4 on XCD (i / 4) % 8: the j-th piece of XCD x
  // takes slot 4 (8 (j / 4) + x) + j % 4; empty slots are padding (frame -1)
  const size_t rounds = (per_max + kWavesPerBlock - 1) / kWavesPerBlock;
  const size_t nslots = rounds * kXcd * kWavesPerBlock;
  std::vector<int4> items(2 * nslots, make_int4(0, 0, -1, 0));
  for (int x = 0; x < kXcd; ++x)
    for (size_t j = 0; j < per_xcd[x].size(); ++j) {
      const size_t slot = kWavesPerBlock * (kXcd * (j / kWavesPerBlock) + x) + j % kWavesPerBlock;
      items[2 * slot] = per_xcd[x][j].a;
      items[2 * slot + 1] = per_xcd[x][j].b;
    }
  const int64_t nslots_used = (int64_t)nslots;
  // axis-0 window values rotated to the slot (cell mod taps) each tap lands on
  std::vector<float> rot((size_t)M * taps);
  for (int64_t s = 0; s < M; ++s) {
    const int b0 = hbase[s * 2];
    for (int k = 0; k < taps; ++k) rot[s * taps + k] = hpsi[(s * 2) * taps + (((k - b0) % taps) + taps) % taps];
  }
  // kSweepPad padding records (sample 0, zero weights) after the last entry:
  // the kernel loads whole chunks of records and skips those past an item's end
  ent.resize(ent.size() + kSweepPad, make_int4(0, 0, 0, 0));
  CHK(dalloc(&p.sw_items, items.size()));
  CHK(dalloc(&p.sw_ent, ent.size()));
  CHK(dalloc(&p.psi_rot, rot.size()));
  HIPCHK(hipMemcpy(p.sw_items, items.data(), sizeof(int4) * items.size(), hipMemcpyHostToDevice));
  if (!ent.empty()) HIPCHK(hipMemcpy(p.sw_ent, ent.data(), sizeof(int4) * ent.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(p.psi_rot, rot.data(), sizeof(float) * rot.size(), hipMemcpyHostToDevice));
  if (!ritems.empty()) {
    CHK(dalloc(&p.sw_ritems, ritems.size()));
    CHK(dalloc(&p.sw_rpieces, rpieces.size()));
    CHK(dalloc(&p.sw_part, (size_t)part_floats));
    HIPCHK(hipMemcpy(p.sw_ritems, ritems.data(), sizeof(int4) * ritems.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(p.sw_rpieces, rpieces.data(), sizeof(int4) * rpieces.size(), hipMemcpyHostToDevice));
  }
  p.n_sw_ritems = (int)ritems.size();
  p.n_sw_items = (int)nslots_used;
  p.n_sw_ent = nent;
  p.sw_nseg = nseg;
  p.h_sw_ent = std::move(ent);
  return CGS_OK;
}

// the weighted field of the line-sweep entries: phi1 * w of the entry's sample
static int sweep_weights(Plan& p, const std::vector<float>& ws) {
  if (!p.sw_ent || p.h_sw_ent.empty()) return CGS_OK;
  for (auto& e : p.h_sw_ent) {
    float w1;
    std::memcpy(&w1, &e.z, 4);
    const float wf = w1 * ws[(size_t)e.x];
    std::memcpy(&e.w, &wf, 4);
  }
  HIPCHK(hipMemcpy(p.sw_ent, p.h_sw_ent.data(), sizeof(int4) * p.h_sw_ent.size(), hipMemcpyHostToDevice));
  return CGS_OK;
}

// sweep: also (re)build the line-sweep tables, which depend on the trajectory
// only (not on the visit-list chunking the autotuner varies)
static int build_samples(Plan& p, const float* k, bool sweep = true) {
  const int d = p.d, m = p.m, taps = p.taps, T = p.T;
  const int64_t M = p.Mtot;
  // base cell and fractional offset of every sample (FP64)
  std::vector<int> cell((size_t)M * d);
  std::vector<double> frac((size_t)M * d);
  std::vector<int> frame((size_t)M);
  for (int f = 0; f < p.F; ++f)
    for (int64_t j = p.foff[f]; j < p.foff[f + 1]; ++j) frame[j] = f;
  for (int64_t j = 0; j < M; ++j) {
    for (int a = 0; a < d; ++a) {
      const double x = (double)k[(d - 1 - a) + (int64_t)d * j];   // k(t,j) -> image axis d-1-t
      if (!(x >= -0.5 && x < 0.5)) {
        set_error("trajectory value outside [-0.5, 0.5) at sample " + std::to_string(j));
        return CGS_ERR_ARG;
      }
      const double u = p.n[a] * x;
      const double fl = std::floor(u);
      int c = (int)((long)fl % p.n[a]);
      if (c < 0) c += p.n[a];
      cell[j * d + a] = c;
      frac[j * d + a] = u - fl;
    }
  }
  // counting sort by (frame, tile of the base cell)
  const int64_t tpf = p.tiles_per_frame;
  const int64_t nkeys = tpf * p.F;
  std::vector<int64_t> key((size_t)M);
  std::vector<int64_t> cnt((size_t)nkeys + 1, 0);
  for (int64_t j = 0; j < M; ++j) {
    int64_t t = 0, mul = 1;
    for (int a = 0; a < d; ++a) { t += (int64_t)(cell[j * d + a] / T) * mul; mul *= p.nt[a]; }
    key[j] = frame[j] * tpf + t;
    cnt[key[j] + 1]++;
  }
  for (int64_t q = 0; q < nkeys; ++q) cnt[q + 1] += cnt[q];
  std::vector<int> perm((size_t)M);
  {
    std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
    for (int64_t j = 0; j < M; ++j) perm[pos[key[j]]++] = (int)j;
  }
  // sorted base (cell - m mod n) and window values
  std::vector<int> hbase((size_t)M * d);
  std::vector<float> hpsi((size_t)M * d * taps);
  std::vector<int> hframe((size_t)M);
  {
    const int nthr = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (int t = 0; t < nthr; ++t) {
      th.emplace_back([&, t]() {
        for (int64_t s = t; s < M; s += nthr) {
          const int64_t j = perm[s];
          hframe[s] = frame[j];
          for (int a = 0; a < d; ++a) {
            int b = cell[j * d + a] - m;
            b %= p.n[a];
            if (b < 0) b += p.n[a];
            hbase[s * d + a] = b;
            for (int q = 0; q < taps; ++q)
              hpsi[(s * d + a) * taps + q] = (float)kb_phi(frac[j * d + a] + m - q, m, p.b[a]);
          }
        }
      });
    }
    for (auto& x : th) x.join();
  }
  // interp items: runs of samples of one frame
  const int ipw = (d == 3) ? 4 : 16;
  std::vector<int4> items;
  for (int f = 0; f < p.F; ++f)
    for (int64_t s = p.foff[f]; s < p.foff[f + 1]; s += ipw)
      items.push_back(make_int4(f, (int)s, (int)std::min<int64_t>(ipw, p.foff[f + 1] - s), 0));
  // tile-level interp items {frame, tile, sbeg, send}: the samples whose base
  // cell lies in one tile, at most tile_item_max per item (heaviest first)
  std::vector<int4> titems;
  for (int64_t kk = 0; kk < nkeys; ++kk) {
    const int f = (int)(kk / tpf), t = (int)(kk % tpf);
    for (int64_t s = cnt[kk]; s < cnt[kk + 1]; s += p.tile_item_max)
      titems.push_back(make_int4(f, t, (int)s, (int)std::min<int64_t>(s + p.tile_item_max, cnt[kk + 1])));
  }
  std::stable_sort(titems.begin(), titems.end(), [](const int4& x, const int4& y) { return (x.w - x.z) > (y.w - y.z); });
  // gridding visit lists: every (tile, sample) pair whose footprint meets the tile
  std::vector<int64_t> vcnt((size_t)nkeys + 1, 0);
  auto tiles_of = [&](int64_t s, int a, int* out) {
    int nout = 0;
    for (int q = 0; q < taps; ++q) {
      int c = hbase[s * d + a] + q;
      if (c >= p.n[a]) c -= p.n[a];
      const int t = c / T;
      bool seen = false;
      for (int z = 0; z < nout; ++z) seen |= (out[z] == t);
      if (!seen) out[nout++] = t;
    }
    return nout;
  };
  auto for_visits = [&](auto&& fn) {
    int ta[3][16], na[3] = {1, 1, 1};
    ta[1][0] = ta[2][0] = 0;
    for (int64_t s = 0; s < M; ++s) {
      for (int a = 0; a < d; ++a) na[a] = tiles_of(s, a, ta[a]);
      const int64_t fbase = (int64_t)hframe[s] * tpf;
      for (int z2 = 0; z2 < na[2]; ++z2)
        for (int z1 = 0; z1 < na[1]; ++z1)
          for (int z0 = 0; z0 < na[0]; ++z0) {
            const int64_t t = ta[0][z0] + (int64_t)p.nt[0] * (ta[1][z1] + (int64_t)p.nt[1] * ta[2][z2]);
            fn(fbase + t, s);
          }
    }
  };
  for_visits([&](int64_t kk, int64_t) { vcnt[kk + 1]++; });
  for (int64_t q = 0; q < nkeys; ++q) vcnt[q + 1] += vcnt[q];
  const int64_t nvis = vcnt[nkeys];
  if (nvis > INT32_MAX) {
    set_error("too many gridding visits for 32-bit indices");
    return CGS_ERR_UNSUPPORTED;
  }
  std::vector<int> vis((size_t)nvis);
  {
    std::vector<int64_t> pos(vcnt.begin(), vcnt.end() - 1);
    for_visits([&](int64_t kk, int64_t s) { vis[pos[kk]++] = (int)s; });
  }
  // visit records: sample id, offsets of the footprint in the tile, in-tile tap ranges
  std::vector<int4> vrec((size_t)nvis);
  for (int64_t kk = 0; kk < nkeys; ++kk) {
    const int t = (int)(kk % tpf);
    const int tt[3] = {t % p.nt[0], (t / p.nt[0]) % p.nt[1], t / (p.nt[0] * p.nt[1])};
    int ts[3], te[3];
    for (int a = 0; a < 3; ++a) {
      ts[a] = tt[a] * T;
      te[a] = (a < d) ? std::min(T, p.n[a] - ts[a]) : 1;
    }
    for (int64_t v = vcnt[kk]; v < vcnt[kk + 1]; ++v) {
      const int s = vis[v];
      int rel[3] = {0, 0, 0}, lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
      for (int a = 0; a < d; ++a) {
        int r = hbase[(int64_t)s * d + a] - ts[a];
        if (r < 0) r += p.n[a];
        if (r >= te[a]) r -= p.n[a];
        rel[a] = r;
        lo[a] = std::max(0, -r);
        hi[a] = std::min(taps, te[a] - r);
      }
      int4 rc;
      rc.x = s;
      rc.y = (int)(((uint32_t)rel[0] & 0xffffu) | ((uint32_t)rel[1] << 16));
      rc.z = (int)((uint32_t)lo[0] | ((uint32_t)hi[0] << 8) | ((uint32_t)lo[1] << 16) | ((uint32_t)hi[1] << 24));
      rc.w = (d == 3) ? (int)(((uint32_t)rel[2] & 0xffffu) | ((uint32_t)lo[2] << 16) | ((uint32_t)hi[2] << 24)) : 0;
      vrec[v] = rc;
    }
  }
  std::vector<GridItem> gitems;
  std::vector<ReduceItem> ritems;
  int slots = 0;
  const int K = p.visits_per_item;
  for (int64_t kk = 0; kk < nkeys; ++kk) {
    const int f = (int)(kk / tpf), t = (int)(kk % tpf);
    const int64_t b0 = vcnt[kk], b1 = vcnt[kk + 1], nv = b1 - b0;
    if (nv <= K) {
      gitems.push_back(GridItem{f, t, (int)b0, (int)b1, -1, 0, 0, 0});
    } else {
      const int nch = (int)((nv + K - 1) / K);
      ritems.push_back(ReduceItem{f, t, slots, nch});
      for (int c = 0; c < nch; ++c) {
        const int64_t c0 = b0 + (int64_t)c * nv / nch, c1 = b0 + (int64_t)(c + 1) * nv / nch;
        gitems.push_back(GridItem{f, t, (int)c0, (int)c1, slots + c, 0, 0, 0});
      }
      slots += nch;
    }
  }
  // longest items first
  std::stable_sort(gitems.begin(), gitems.end(),
                   [](const GridItem& x, const GridItem& y) { return (x.vend - x.vbeg) > (y.vend - y.vbeg); });

  free_samples(p, sweep);
  CHK(dalloc(&p.base, hbase.size()));
  CHK(dalloc(&p.psi, hpsi.size()));
  CHK(dalloc(&p.perm, perm.size()));
  CHK(dalloc(&p.iperm, perm.size()));
  CHK(dalloc(&p.sframe, hframe.size()));
  CHK(dalloc(&p.interp_items, items.size()));
  CHK(dalloc(&p.tile_items, titems.size()));
  if (!titems.empty())
    HIPCHK(hipMemcpy(p.tile_items, titems.data(), sizeof(int4) * titems.size(), hipMemcpyHostToDevice));
  p.n_tile_items = (int)titems.size();
  CHK(dalloc(&p.grid_items, gitems.size()));
  CHK(dalloc(&p.reduce_items, ritems.size()));
  CHK(dalloc(&p.visits, vis.size()));
  CHK(dalloc(&p.vrec, vrec.size()));
  if (!vrec.empty()) HIPCHK(hipMemcpy(p.vrec, vrec.data(), sizeof(int4) * vrec.size(), hipMemcpyHostToDevice));
  const int64_t tilecells = (int64_t)T * T * (d == 3 ? T : 1);
  CHK(dalloc(&p.partial, (size_t)slots * tilecells * p.C));
  HIPCHK(hipMemcpy(p.base, hbase.data(), sizeof(int) * hbase.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(p.psi, hpsi.data(), sizeof(float) * hpsi.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(p.perm, perm.data(), sizeof(int) * perm.size(), hipMemcpyHostToDevice));
  {
    std::vector<int> iperm(perm.size());
    for (size_t s = 0; s < perm.size(); ++s) iperm[(size_t)perm[s]] = (int)s;
    HIPCHK(hipMemcpy(p.iperm, iperm.data(), sizeof(int) * iperm.size(), hipMemcpyHostToDevice));
  }
  HIPCHK(hipMemcpy(p.sframe, hframe.data(), sizeof(int) * hframe.size(), hipMemcpyHostToDevice));
  if (!items.empty())
    HIPCHK(hipMemcpy(p.interp_items, items.data(), sizeof(int4) * items.size(), hipMemcpyHostToDevice));
  if (!gitems.empty())
    HIPCHK(hipMemcpy(p.grid_items, gitems.data(), sizeof(GridItem) * gitems.size(), hipMemcpyHostToDevice));
  if (!ritems.empty())
    HIPCHK(hipMemcpy(p.reduce_items, ritems.data(), sizeof(ReduceItem) * ritems.size(), hipMemcpyHostToDevice));
  if (!vis.empty()) HIPCHK(hipMemcpy(p.visits, vis.data(), sizeof(int) * vis.size(), hipMemcpyHostToDevice));
  if (sweep) CHK(build_sweep(p, hbase, hpsi, hframe));
  if (sweep) CHK(build_bins(p, hbase, hpsi, hframe));
  p.n_interp_items = (int)items.size();
  p.n_grid_items = (int)gitems.size();
  p.n_reduce_items = (int)ritems.size();
  p.n_visits = nvis;
  p.n_slots = slots;
  p.h_perm = std::move(perm);
  return CGS_OK;
}

// Binned gridding tables (kernels_bin.hip), built only when the binned scatter
// is switched on (CGS_GRID_BIN=1): on gfx950 its LDS float atomics take ~128
// cycles per 64-lane ds_add_f32 (LDS 80-90% busy, no bank conflicts), which
// makes it 30-40x slower than the tile kernels.  The bin side TB (2D: 32, else 16;
// 3D: 16, else 8) must divide every grid axis and leave the region, TB + taps - 1
// cells, no longer than the axis; the region of one coil group must fit in LDS.
// Otherwise (and with CGS_DETERMINISTIC) the plan has no bins.  Items hold at
// most bin_item_max samples (2D 2048, 3D 32768; a heavy bin's chunks are equal)
// and run longest first; each owns a region slot, a bin's slots consecutive.
static int build_bins(Plan& p, const std::vector<int>& hbase, const std::vector<float>& hpsi,
                      const std::vector<int>& hframe) {
  free_bins(p);
  if (!p.bin_allowed || (p.flags & CGS_DETERMINISTIC)) return CGS_OK;
  const int d = p.d, taps = p.taps;
  std::vector<int> cands = (d == 2) ? std::vector<int>{32, 16} : std::vector<int>{16, 8};
  if (const char* e = std::getenv("CGS_BIN_TB")) cands = {std::atoi(e)};
  int CG = p.bin_cg;
  if (d == 2 && CG != 4 && CG != 8) CG = 8;
  if (d == 3 && CG != 1 && CG != 2) CG = 1;
  int TB = 0;
  for (int tb : cands) {
    if (tb < 8 || tb > 64) continue;
    bool ok = grid_bin_lds(d, taps, CG, tb + taps - 1, bin_row_stride(tb + taps - 1)) <= 160 * 1024;
    for (int a = 0; a < d; ++a) ok = ok && p.n[a] % tb == 0 && p.n[a] >= tb + taps - 1;
    if (ok) { TB = tb; break; }
  }
  if (!TB) return CGS_OK;
  const int E = TB + taps - 1;
  int nb[3] = {1, 1, 1};
  int64_t nbins = 1, ecells = 1;
  for (int a = 0; a < d; ++a) { nb[a] = p.n[a] / TB; nbins *= nb[a]; ecells *= E; }
  const int64_t M = p.Mtot, nkeys = nbins * p.F;
  if (nkeys > INT32_MAX) return CGS_OK;
  std::vector<int64_t> cnt((size_t)nkeys + 1, 0);
  std::vector<int64_t> key((size_t)M);
  for (int64_t s = 0; s < M; ++s) {
    int64_t b = 0, mul = 1;
    for (int a = 0; a < d; ++a) { b += (int64_t)(hbase[s * d + a] / TB) * mul; mul *= nb[a]; }
    key[s] = (int64_t)hframe[s] * nbins + b;
    cnt[key[s] + 1]++;
  }
  for (int64_t q = 0; q < nkeys; ++q) cnt[q + 1] += cnt[q];
  std::vector<int> ids((size_t)M);
  {
    std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
    for (int64_t s = 0; s < M; ++s) ids[pos[key[s]]++] = (int)s;
  }
  // per id: packed offsets of the footprint in its bin's region, then the window values
  const int R = 1 + d * taps;
  std::vector<float> rec((size_t)M * R);
  for (int64_t k = 0; k < M; ++k) {
    const int64_t s = ids[k];
    uint32_t off = 0;
    for (int a = 0; a < d; ++a) off |= (uint32_t)(hbase[s * d + a] % TB) << (8 * a);
    std::memcpy(&rec[k * R], &off, 4);
    for (int q = 0; q < d * taps; ++q) rec[k * R + 1 + q] = hpsi[s * d * taps + q];
  }
  const int K = p.bin_item_max > 0 ? p.bin_item_max : (d == 2 ? 2048 : 32768);
  std::vector<int4> items;
  std::vector<int2> range((size_t)nkeys, make_int2(0, 0));
  int slot = 0;
  for (int64_t kk = 0; kk < nkeys; ++kk) {
    const int64_t b0 = cnt[kk], n = cnt[kk + 1] - b0;
    if (n == 0) continue;
    const int nch = (int)((n + K - 1) / K);
    range[kk] = make_int2(slot, nch);
    for (int c = 0; c < nch; ++c)
      items.push_back(make_int4((int)kk, (int)(b0 + n * c / nch), (int)(b0 + n * (c + 1) / nch), slot + c));
    slot += nch;
  }
  std::stable_sort(items.begin(), items.end(), [](const int4& x, const int4& y) { return (x.z - x.y) > (y.z - y.y); });
  const int ngroups = (p.C + CG - 1) / CG;
  const size_t regn = (size_t)slot * ngroups * (size_t)ecells * CG;
  if (regn * sizeof(float2) > ((size_t)16 << 30)) return CGS_OK;   // region buffer above 16 GB: no bins
  CHK(dalloc(&p.bin_ids, ids.size()));
  CHK(dalloc(&p.bin_items, std::max<size_t>(1, items.size())));
  CHK(dalloc(&p.bin_range, range.size()));
  CHK(dalloc(&p.bin_rec, rec.size()));
  CHK(dalloc(&p.bin_reg, std::max<size_t>(1, regn)));
  HIPCHK(hipMemcpy(p.bin_ids, ids.data(), sizeof(int) * ids.size(), hipMemcpyHostToDevice));
  if (!items.empty()) HIPCHK(hipMemcpy(p.bin_items, items.data(), sizeof(int4) * items.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(p.bin_range, range.data(), sizeof(int2) * range.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(p.bin_rec, rec.data(), sizeof(float) * rec.size(), hipMemcpyHostToDevice));
  p.bin_TB = TB;
  p.bin_cg = CG;
  p.bin_nbins = nbins;
  for (int a = 0; a < 3; ++a) p.bin_nb[a] = nb[a];
  p.n_bin_items = (int)items.size();
  return CGS_OK;
}

static void bin_args(const Plan& p, const float2* y, bool weighted, const int* fl, BinArgs& a) {
  a = BinArgs{};
  a.ids = p.bin_ids; a.items = p.bin_items; a.range = p.bin_range; a.brec = p.bin_rec;
  a.w = weighted ? p.w : nullptr; a.y = y; a.reg = p.bin_reg; a.grid = p.grid; a.fdone = fl;
  a.d = p.d; a.taps = p.taps; a.C = p.C; a.CG = p.bin_cg; a.ngroups = (p.C + p.bin_cg - 1) / p.bin_cg;
  a.TB = p.bin_TB; a.E = p.bin_TB + p.taps - 1; a.P1 = bin_row_stride(a.E);
  a.n0 = p.n[0]; a.n1 = p.n[1]; a.n2 = p.n[2];
  a.nb0 = p.bin_nb[0]; a.nb1 = p.bin_nb[1]; a.nb2 = p.bin_nb[2];
  a.F = p.F; a.nitems = p.n_bin_items;
  a.ntot = p.ntot;
  a.ecells = 1;
  for (int k = 0; k < p.d; ++k) a.ecells *= a.E;
  a.nbins = p.bin_nbins;
}

// ---- off-resonance (B0) model ------------------------------------------------------------
// f_j = sum_r x(r) exp(-2 pi i k_j.r) exp(-2 pi i b0(r) t_j) with the time-frequency
// term as a third NFFT dimension (the mri_inh_3d plan codeare's NFFT builds for
// "b0" and "timing", NFFT.hpp:117-173, NFFTTraits.hpp:91-108): kernels_b0.hip.
constexpr double kB0Sigma = 1.2;   // oversampling of the third axis (NFFT.hpp:147)

static void b0_free(Plan& p) {
  if (p.b0q) {
    destroy_plan(*p.b0q);
    delete p.b0q;
    p.b0q = nullptr;
  }
  dfree(p.b0_phase); dfree(p.b0_nu); dfree(p.b0_phinv);
  p.b0_N3 = 0;
}

// weights of the helper plan in its sample order
static int b0_weights(Plan& p) {
  if (!p.b0q || p.h_w.empty()) return CGS_OK;
  Plan& q = *p.b0q;
  std::vector<float> ws((size_t)q.Mtot);
  for (int64_t s = 0; s < q.Mtot; ++s) ws[s] = p.h_w[q.h_perm[s]];
  HIPCHK(hipMemcpy(q.w, ws.data(), sizeof(float) * ws.size(), hipMemcpyHostToDevice));
  q.h_w = p.h_w;
  q.have_w = true;
  return CGS_OK;
}

// (re)builds the helper plan and the tables from p.h_b0, p.h_t and the trajectory
static int b0_build(Plan& p) {
  b0_free(p);
  if (p.h_b0.empty()) return CGS_OK;
  const int m = p.m;
  double tmin = p.h_t[0], tmax = p.h_t[0], B = 0.0;
  for (float v : p.h_t) { tmin = std::min(tmin, (double)v); tmax = std::max(tmax, (double)v); }
  for (float v : p.h_b0) B = std::max(B, std::fabs((double)v));
  const double ts = 0.5 * (tmin + tmax), T = 0.5 * (tmax - tmin);
  // third axis: N3 >= 4 sigma3 T B + 2m keeps |tau| <= 1/(2 sigma3) with every
  // window tap inside [-N3/2, N3/2) (the size rule of NFFT.hpp:150-155); even,
  // with an oversampled length the FFT passes support and two windows wide
  int N3 = (int)std::ceil(4.0 * kB0Sigma * T * B + 2.0 * m);
  N3 = std::max(N3 + (N3 & 1), 2 * m + 2);
  for (;; N3 += 2) {
    const int n3 = (int)std::ceil(p.sigma * N3);
    int rad[8], nst = 0;
    if (n3 >= 2 * p.taps && factorize(n3, rad, &nst)) break;
    if (N3 > 4096) { set_error("off-resonance axis too long (b0 x timing span)"); return CGS_ERR_UNSUPPORTED; }
  }
  const double omega = (B > 0.0) ? B / (0.5 - (double)m / N3) : 1.0;
  const double b3 = M_PI * (2.0 - 1.0 / kB0Sigma);
  Plan* q = new Plan;
  const int N3d[3] = {p.N[0], p.N[1], N3};
  int rc = create_plan(*q, 3, N3d, p.Mtot, p.sigma, m, p.C, 1, nullptr, CGS_DEVICE_IO);
  if (rc != CGS_OK) { destroy_plan(*q); delete q; return rc; }
  (void)hipStreamDestroy(q->stream);   // the helper runs on the plan's stream
  q->stream = p.stream;
  q->own_stream = false;
  p.b0q = q;
  // nodes (tau_j, k(0, j), k(1, j)): tau pairs with the third image axis
  std::vector<float> k3((size_t)3 * p.Mtot);
  std::vector<double> tau((size_t)p.Mtot);
  for (int64_t j = 0; j < p.Mtot; ++j) {
    tau[j] = ((double)p.h_t[j] - ts) * omega / N3;
    k3[3 * j + 0] = (float)tau[j];
    k3[3 * j + 1] = p.h_k[2 * j + 0];
    k3[3 * j + 2] = p.h_k[2 * j + 1];
  }
  CHK(build_samples(*q, k3.data(), false));
  q->h_k = k3;
  q->have_k = true;
  CHK(b0_weights(p));
  std::vector<float> phinv((size_t)p.Mtot);
  for (int64_t s = 0; s < p.Mtot; ++s) {
    const double a = 2.0 * M_PI * tau[q->h_perm[s]];
    const double s2 = b3 * b3 - a * a;
    phinv[s] = (float)(1.0 / bessel_i0(m * std::sqrt(s2 > 0 ? s2 : 0.0)));
  }
  std::vector<float2> ph((size_t)p.Ntot);
  std::vector<float> nu((size_t)p.Ntot);
  for (int64_t i = 0; i < p.Ntot; ++i) {
    const double b = p.h_b0[i];
    ph[i] = make_float2((float)std::cos(-2.0 * M_PI * b * ts), (float)std::sin(-2.0 * M_PI * b * ts));
    nu[i] = (float)(N3 * b / omega);
  }
  CHK(dalloc(&p.b0_phase, ph.size()));
  CHK(dalloc(&p.b0_nu, nu.size()));
  CHK(dalloc(&p.b0_phinv, phinv.size()));
  HIPCHK(hipMemcpy(p.b0_phase, ph.data(), sizeof(float2) * ph.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(p.b0_nu, nu.data(), sizeof(float) * nu.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(p.b0_phinv, phinv.data(), sizeof(float) * phinv.size(), hipMemcpyHostToDevice));
  p.b0_N3 = N3;
  p.b0_b3 = (float)b3;
  p.b0_ts = ts;
  p.b0_omega = omega;
  return CGS_OK;
}

// y (helper's sample order) = B0-NFFT of the coil images X (p's coil layout),
// with the damping mask on X when masked
static int b0_forward(Plan& p, const float2* X, float2* y, bool masked) {
  Plan& q = *p.b0q;
  launch_b0_spread(X, p.b0_phase, p.b0_nu, q.Xc, p.Ntot, p.C, p.b0_N3, p.m, p.b0_b3, masked ? p.N[0] / 2 : 0,
                   p.N[0], p.N[1], p.stream);
  CHK(fft_forward(q, 2, q.Xc, nullptr));
  CHK(interp(q, y, nullptr));
  launch_b0_scale(y, p.b0_phinv, y, q.Mtot, q.C, p.stream);
  return CGS_OK;
}

// X (p's coil layout) = adjoint B0-NFFT of y (helper's order), optionally of w .* y
static int b0_adjoint(Plan& p, const float2* y, bool weighted, float2* X) {
  Plan& q = *p.b0q;
  launch_b0_scale(y, p.b0_phinv, q.ybuf[2], q.Mtot, q.C, p.stream);
  CHK(gridding(q, q.ybuf[2], weighted, nullptr));
  CHK(fft_adjoint(q, q.Xc, nullptr));
  launch_b0_gather(q.Xc, p.b0_phase, p.b0_nu, X, p.Ntot, p.C, p.b0_N3, p.m, p.b0_b3, p.stream);
  return CGS_OK;
}

// ---- plan life cycle --------------------------------------------------------------------
static void destroy_plan(Plan& p) {
  if (p.stream) (void)hipStreamSynchronize(p.stream);
  b0_free(p);
  free_samples(p);
  for (int a = 0; a < 3; ++a) { dfree(p.fa[a].tw); dfree(p.ft[a].tw); dfree(p.dinv[a]); }
  toep_drop(p);
  dfree(p.d_foff); dfree(p.S); dfree(p.Imap); dfree(p.w);
  dfree(p.grid); dfree(p.t1); dfree(p.t2); dfree(p.Xc); dfree(p.Xc2); dfree(p.Xc3);
  for (auto& b : p.ybuf) dfree(b);
  for (auto& b : p.img) dfree(b);
  dfree(p.ext_a); dfree(p.ext_b); dfree(p.partials); dfree(p.cgst); dfree(p.cnr);
  dfree(p.fdone); dfree(p.fmask); dfree(p.hist); dfree(p.tickets);
  if (p.h_cgst) (void)hipHostFree(p.h_cgst);
  if (p.h_done) (void)hipHostFree(p.h_done);
  p.h_cgst = nullptr; p.h_done = nullptr;
  for (auto& q : p.pending) { p.event_pool.push_back(q.a); p.event_pool.push_back(q.b); }
  p.pending.clear();
  if (p.graph_exec) (void)hipGraphExecDestroy(p.graph_exec);
  p.graph_exec = nullptr;
  for (auto& e : p.poll_ev) {
    if (e) (void)hipEventDestroy(e);
    e = nullptr;
  }
  for (auto e : p.event_pool) (void)hipEventDestroy(e);
  p.event_pool.clear();
  if (p.own_stream && p.stream) (void)hipStreamDestroy(p.stream);
  p.stream = nullptr;
}

static int create_plan(Plan& p, int d, const int* N, int64_t Mtot, double sigma, int m, int C, int F,
                       const int64_t* foff, unsigned flags) {
  if (d != 2 && d != 3) { set_error("d must be 2 or 3"); return CGS_ERR_UNSUPPORTED; }
  if (!N || Mtot < 1 || C < 1 || F < 1 || !(sigma >= 1.0)) { set_error("bad plan arguments"); return CGS_ERR_ARG; }
  if (m < 1 || m > 4) { set_error("kernel cut-off m must be 1..4"); return CGS_ERR_UNSUPPORTED; }
  if (C > 512) { set_error("at most 512 coils"); return CGS_ERR_UNSUPPORTED; }
  p.d = d; p.m = m; p.taps = 2 * m + 2; p.C = C; p.F = F; p.Mtot = Mtot; p.flags = flags;
  p.Ntot = 1; p.ntot = 1;
  for (int a = 0; a < d; ++a) {
    if (N[a] < 2 || (N[a] & 1)) { set_error("image sides must be even and >= 2"); return CGS_ERR_ARG; }
    p.N[a] = N[a];
    p.n[a] = (int)std::ceil(sigma * (double)N[a]);   // NFFT.hpp:108-109
    {
      int rtmp[8], ntmp = 0;
      if (!factorize(p.n[a], rtmp, &ntmp)) {
        set_error("oversampled size " + std::to_string(p.n[a]) + " is not 2^a 3^b");
        return CGS_ERR_UNSUPPORTED;
      }
    }
    if (p.n[a] < 2 * p.taps) { set_error("oversampled grid smaller than two window widths"); return CGS_ERR_UNSUPPORTED; }
    if (p.n[a] > 4096) { set_error("oversampled side above 4096"); return CGS_ERR_UNSUPPORTED; }
    p.Ntot *= N[a];
    p.ntot *= p.n[a];
    const double s = (double)p.n[a] / (double)N[a];
    p.b[a] = M_PI * (2.0 - 1.0 / s);
  }
  p.foff.assign(F + 1, 0);
  if (F == 1) { p.foff[1] = Mtot; }
  else {
    if (!foff) { set_error("frame_offsets required for several frames"); return CGS_ERR_ARG; }
    for (int f = 0; f <= F; ++f) p.foff[f] = foff[f];
    if (p.foff[0] != 0 || p.foff[F] != Mtot) { set_error("frame_offsets must run from 0 to M_total"); return CGS_ERR_ARG; }
    for (int f = 0; f < F; ++f)
      if (p.foff[f + 1] < p.foff[f]) { set_error("frame_offsets must not decrease"); return CGS_ERR_ARG; }
  }
  if (Mtot > INT32_MAX / 2) { set_error("too many samples"); return CGS_ERR_UNSUPPORTED; }
  // tiles and kernel shapes
  p.T = (d == 3) ? 8 : 8;
  for (int a = 0; a < 3; ++a) p.nt[a] = (a < d) ? (p.n[a] + p.T - 1) / p.T : 1;
  p.tiles_per_frame = (int64_t)p.nt[0] * p.nt[1] * p.nt[2];
  // batched gridding: CG coils per block, W = CG/8 waves so that the W*RS = 8
  // row slots divide the tile side
  p.use_grid_batched = true;
  p.grid_cg = (C <= 8) ? 8 : (C <= 16 ? 16 : 32);
  p.grid_w = p.grid_cg / 8;
  p.interp_cg = (C <= 8) ? 8 : 16;
  // 3D: the LDS-staged gather stages the (8 + taps - 1)^3 x 2-coil region once
  // per item and streams the item's samples through it 64 at a time; items of
  // up to 2048 samples amortise the staging (dense koosh-ball tiles hold
  // thousands of samples)
  p.use_interp_lds = true;
  // 2D: (coil pair, sample) in 512-thread blocks (C2's gather: 62.9 us against
  // 72-74 us for one coil a lane); 3D: (sample, plane, tap parity) where taps = 8, T = 8
  p.interp_lanes = (d == 2) ? 4 : 2;
  p.tile_item_max = (d == 3) ? 2048 : 128;
  p.visits_per_item = (d == 3) ? 4096 : 512;
  p.grid_cgr = (C <= 8) ? 8 : 16;
  // shape overrides for tuning sweeps (tile side must be a multiple of 8)
  if (const char* e = std::getenv("CGS_TILE")) {
    const int t = std::atoi(e);
    if (t >= 8 && t % 8 == 0) {
      p.T = t;
      for (int a = 0; a < 3; ++a) p.nt[a] = (a < d) ? (p.n[a] + p.T - 1) / p.T : 1;
      p.tiles_per_frame = (int64_t)p.nt[0] * p.nt[1] * p.nt[2];
    }
  }
  if (const char* e = std::getenv("CGS_GRID_CG")) {
    const int g = std::atoi(e);
    if (g == 8 || g == 16 || g == 32) { p.grid_cg = g; p.grid_w = g / 8; }
  }
  if (const char* e = std::getenv("CGS_INTERP_CG")) {
    const int g = std::atoi(e);
    if (g == 8 || g == 16 || g == 32) p.interp_cg = g;
  }
  if (const char* e = std::getenv("CGS_VISITS_PER_ITEM")) p.visits_per_item = std::max(16, std::atoi(e));
  if (const char* e = std::getenv("CGS_TILE_ITEM_MAX")) p.tile_item_max = std::max(8, std::atoi(e));
  if (const char* e = std::getenv("CGS_INTERP_LDS")) p.use_interp_lds = std::atoi(e) != 0;
  if (const char* e = std::getenv("CGS_INTERP_CS")) p.interp_lanes = std::max(0, std::min(4, std::atoi(e)));
  if (const char* e = std::getenv("CGS_INTERP3_THREADS")) p.interp3_threads = std::atoi(e);
  if (const char* e = std::getenv("CGS_GRID_REG")) p.use_grid_reg = std::atoi(e) != 0;
  p.use_grid_mfma = true;
  if (const char* e = std::getenv("CGS_GRID_MFMA")) p.use_grid_mfma = std::atoi(e) != 0;
  if (const char* e = std::getenv("CGS_GRID_CGR")) {
    const int g = std::atoi(e);
    if (g == 8 || g == 16 || g == 32) p.grid_cgr = g;
  }
  if (const char* e = std::getenv("CGS_GRID_ATOMIC")) p.use_grid_atomic = std::atoi(e) != 0;
  if (const char* e = std::getenv("CGS_GRID_BATCHED")) p.use_grid_batched = std::atoi(e) != 0;
  if (const char* e = std::getenv("CGS_GRID_BIN")) p.use_grid_bin = p.bin_allowed = std::atoi(e) != 0;
  if (const char* e = std::getenv("CGS_CG_TAILFOLD")) p.tailfold = std::atoi(e) != 0;
  if (const char* e = std::getenv("CGS_BIN_CG")) p.bin_cg = std::atoi(e);
  if (const char* e = std::getenv("CGS_BIN_K")) p.bin_item_max = std::max(64, std::atoi(e));
  if (const char* e = std::getenv("CGS_BIN_THREADS")) p.bin_threads = std::atoi(e);
  if (const char* e = std::getenv("CGS_GRID_SWEEP")) p.use_grid_sweep = p.sweep_allowed = std::atoi(e) != 0;
  if (const char* e = std::getenv("CGS_SWEEP_W")) p.sweep_W = std::max(4, std::atoi(e));
  if (const char* e = std::getenv("CGS_SWEEP_K")) p.sweep_K = std::max(16, std::atoi(e));

  if (flags & 0x8000u) {   // internal: caller-provided stream set later
  }
  HIPCHK(hipGetDevice(&p.device));
  HIPCHK(hipStreamCreateWithFlags(&p.stream, hipStreamNonBlocking));
  p.own_stream = true;
  HIPCHK(hipEventCreateWithFlags(&p.poll_ev[0], hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&p.poll_ev[1], hipEventDisableTiming));
  p.done_ms.assign((size_t)F, -1.0);

  // tables
  p.sigma = sigma;
  for (int a = 0; a < d; ++a) {
    CHK(make_tab(p.fa[a], p.n[a]));
    std::vector<float> di(p.N[a]);
    for (int i = 0; i < p.N[a]; ++i) {
      const double kk = 2.0 * M_PI * (double)(i - p.N[a] / 2) / (double)p.n[a];
      const double s2 = p.b[a] * p.b[a] - kk * kk;
      di[i] = (float)(1.0 / bessel_i0(m * std::sqrt(s2 > 0 ? s2 : 0.0)));   // 1/phi_hat
    }
    CHK(dalloc(&p.dinv[a], di.size()));
    HIPCHK(hipMemcpy(p.dinv[a], di.data(), sizeof(float) * di.size(), hipMemcpyHostToDevice));
  }
  CHK(dalloc(&p.d_foff, (size_t)F + 1));
  HIPCHK(hipMemcpy(p.d_foff, p.foff.data(), sizeof(int64_t) * (F + 1), hipMemcpyHostToDevice));

  // buffers
  const size_t FC = (size_t)F * C;
  CHK(dalloc(&p.grid, FC * p.ntot));
  size_t t1n = (size_t)C * p.n[0] * p.N[1] * (d == 3 ? p.N[2] : 1) * F;
  CHK(dalloc(&p.t1, t1n));
  if (d == 3) CHK(dalloc(&p.t2, (size_t)C * p.n[0] * p.n[1] * p.N[2] * F));
  CHK(dalloc(&p.Xc, FC * p.Ntot));
  CHK(dalloc(&p.Xc2, FC * p.Ntot));
  CHK(dalloc(&p.Xc3, FC * p.Ntot));
  for (auto& b : p.ybuf) CHK(dalloc(&b, (size_t)Mtot * C));
  for (auto& b : p.img) CHK(dalloc(&b, (size_t)F * p.Ntot));
  CHK(dalloc(&p.w, (size_t)Mtot));
  const int nmaps = (flags & CGS_MAPS_PER_FRAME) ? F : 1;
  CHK(dalloc(&p.S, (size_t)p.Ntot * C * nmaps));
  CHK(dalloc(&p.Imap, (size_t)p.Ntot * nmaps));
  p.Sfs = nmaps > 1 ? p.Ntot * C : 0;
  p.Ifs = nmaps > 1 ? p.Ntot : 0;
  p.n_partials = std::max<size_t>((size_t)4096 * F, (size_t)256 * F * C);
  CHK(dalloc(&p.partials, p.n_partials));
  CHK(dalloc(&p.cgst, (size_t)F));
  CHK(dalloc(&p.tickets, (size_t)2 * F));
  HIPCHK(hipMemset(p.tickets, 0, sizeof(int) * 2 * F));
  CHK(dalloc(&p.cnr, (size_t)8 * F * C));
  CHK(dalloc(&p.fdone, (size_t)F));
  CHK(dalloc(&p.fmask, (size_t)F));
  HIPCHK(hipMemset(p.fdone, 0, sizeof(int) * F));
  HIPCHK(hipMemset(p.fmask, 0, sizeof(int) * F));
  HIPCHK(hipHostMalloc((void**)&p.h_cgst, sizeof(CgState) * F, hipHostMallocDefault));
  HIPCHK(hipHostMalloc((void**)&p.h_done, sizeof(int) * 2 * F, hipHostMallocDefault));
  if (!(flags & CGS_DEVICE_IO)) {
    p.ext_a_bytes = sizeof(float2) * (size_t)Mtot * C;
    p.ext_b_bytes = sizeof(float2) * (size_t)p.Ntot * std::max(F, C);
    CHK(dalloc((float2**)&p.ext_a, (size_t)Mtot * C));
    CHK(dalloc((float2**)&p.ext_b, (size_t)p.Ntot * std::max(F, C)));
  }
  return CGS_OK;
}

static int require_ready(const Plan& p, bool need_S) {
  if (!p.have_k) { set_error("trajectory not set (cgs_set_trajectory)"); return CGS_ERR_STATE; }
  if (need_S && p.b0q) {
    // codeare's NCSENSE reads "b0" but passes it to none of its NFFTs
    // (NCSENSE.hpp:100-103): the coil operators have no off-resonance model
    set_error("the NCSENSE operators do not model off-resonance: clear b0 (cgs_set_b0 with NULL)");
    return CGS_ERR_UNSUPPORTED;
  }
  if (!p.have_w) { set_error("weights not set (cgs_set_weights)"); return CGS_ERR_STATE; }
  if (need_S && !p.have_S) { set_error("sensitivities not set (cgs_set_sensitivities)"); return CGS_ERR_STATE; }
  return CGS_OK;
}

// frame mask with only `frame` active
static void mask_one_frame(Plan& p, int frame) {
  launch_fill_int(p.fmask, 1, p.F, p.stream);
  launch_fill_int(p.fmask + frame, 0, 1, p.stream);
}

// stage host input into a device buffer (or pass a device pointer through)
static int stage_in(Plan& p, const void* src, float2* staging, size_t bytes, const float2** out) {
  if (p.flags & CGS_DEVICE_IO) { *out = (const float2*)src; return CGS_OK; }
  HIPCHK(hipMemcpyAsync(staging, src, bytes, hipMemcpyHostToDevice, p.stream));
  *out = staging;
  return CGS_OK;
}
static int stage_out(Plan& p, void* dst, const float2* dev, size_t bytes) {
  if (p.flags & CGS_DEVICE_IO) return CGS_OK;
  HIPCHK(hipMemcpyAsync(dst, dev, bytes, hipMemcpyDeviceToHost, p.stream));
  HIPCHK(hipStreamSynchronize(p.stream));
  return CGS_OK;
}

static int check_launch() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(std::string("kernel launch failed: ") + hipGetErrorString(e));
    return CGS_ERR_HIP;
  }
  return CGS_OK;
}

// ---- Toeplitz embedding of E^H W E ------------------------------------------------------
// (E^H W E x)(i) = sum_c conj(S_c(i)) sum_i' K(i - i') S_c(i') x(i'),
// K(D) = sum_j w_j exp(+2 pi i k_j . D), D in [-N, N)^d.
// K is the adjoint NFFT of the weights onto a (2N)^d image (a helper plan with
// image side 2N); its (2N)^d FFT divided by (2N)^d is the real spectrum That,
// so the coil convolution becomes zero-pad -> FFT_2N -> .* That -> IFFT_2N -> crop.
// The gridding and the window gather then leave the CG loop entirely.
static int toeplitz_prepare(Plan& p) {
  if (p.have_toep) return CGS_OK;
  if (p.h_k.empty() || p.h_w.empty()) { set_error("Toeplitz operator needs trajectory and weights"); return CGS_ERR_STATE; }
  const int d = p.d, F = p.F;
  int N2[3] = {1, 1, 1};
  int64_t tot2 = 1;
  for (int a = 0; a < d; ++a) {
    N2[a] = 2 * p.N[a];
    tot2 *= N2[a];
    if (!p.ft[a].tw) CHK(make_tab(p.ft[a], N2[a]));
  }
  HIPCHK(hipStreamSynchronize(p.stream));
  Plan q;
  int rc = create_plan(q, d, N2, p.Mtot, 2.0, p.m, 1, F, p.foff.data(), CGS_DEVICE_IO);
  if (rc == CGS_OK) rc = build_samples(q, p.h_k.data());
  if (rc != CGS_OK) { destroy_plan(q); return rc; }
  {
    std::vector<float2> wy((size_t)q.Mtot);
    for (int64_t s = 0; s < q.Mtot; ++s) wy[s] = make_float2(p.h_w[q.h_perm[s]], 0.f);
    hipError_t e = hipMemcpy(q.ybuf[0], wy.data(), sizeof(float2) * wy.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) { destroy_plan(q); set_error("copy of weights failed"); return CGS_ERR_HIP; }
  }
  gridding(q, q.ybuf[0], false, nullptr);   // K(D) per frame -> q.Xc [f][(2N)^d]
  fft_adjoint(q, q.Xc, nullptr);
  // forward (2N)^d FFT of K, with K(D) at bin D mod 2N (the pad mapping with nin = n)
  const float2* src = q.Xc;
  float2* bufs[2] = {q.t1, q.grid};
  int64_t lower = 1, upper = tot2;
  for (int a = 0; a < d; ++a) {
    upper /= N2[a];
    FftArgs fa = fft_base(q, p.ft[a], -1);
    fa.in = src;
    fa.out = bufs[a & 1];
    fa.nin = fa.nout = N2[a];
    fa.inner = lower;
    fa.outer_per_frame = upper;
    fa.outers = upper * F;
    fft_finish(fa);
    launch_fft(fa, 0, 0, q.stream);
    src = fa.out;
    lower *= N2[a];
  }
  rc = CGS_OK;
  {
    const int64_t cells = tot2 / N2[d - 1];
    hipError_t e = hipMalloc((void**)&p.That, sizeof(float) * (size_t)tot2 * F);
    if (e != hipSuccess) { destroy_plan(q); set_error("allocation of the Toeplitz spectrum failed"); return CGS_ERR_ALLOC; }
    launch_toep_layout(src, p.That, cells, N2[d - 1], F, 1.0f / (float)tot2, q.stream);
  }
  {
    const size_t rows = (size_t)(p.Ntot / p.N[0]) * F;
    const size_t groups = (size_t)(p.C + 15) / 16;
    if (hipMalloc((void**)&p.tail_pp, sizeof(double) * rows) != hipSuccess ||
        hipMalloc((void**)&p.tail_pq, sizeof(double) * rows * groups) != hipSuccess ||
        hipMalloc((void**)&p.fold_rq, sizeof(double) * rows) != hipSuccess ||
        hipMalloc((void**)&p.fold_qq, sizeof(double) * rows) != hipSuccess ||
        hipMalloc((void**)&p.fold_rr, sizeof(double) * rows) != hipSuccess) {
      destroy_plan(q);
      set_error("allocation of the CG tail partials failed");
      return CGS_ERR_ALLOC;
    }
  }
  hipError_t e = hipStreamSynchronize(q.stream);
  destroy_plan(q);
  if (e != hipSuccess) { set_error(std::string("Toeplitz preparation failed: ") + hipGetErrorString(e)); return CGS_ERR_HIP; }
  // intermediates: alias the NFFT ones when n == 2N
  bool alias = true;
  for (int a = 0; a < d; ++a) alias &= (p.n[a] == N2[a]);
  if (alias) {
    p.tt1 = p.t1;
    p.tt2 = p.t2;
  } else {
    CHK(dalloc(&p.tt1, (size_t)p.C * N2[0] * p.N[1] * (d == 3 ? p.N[2] : 1) * F));
    if (d == 3) CHK(dalloc(&p.tt2, (size_t)p.C * N2[0] * N2[1] * p.N[2] * F));
  }
  p.tt_alias = alias;
  p.have_toep = true;
  return CGS_OK;
}

// drops the Toeplitz spectrum and any captured CG graph (both refer to the
// trajectory-dependent buffers)
static void toep_drop(Plan& p) {
  dfree(p.That);
  dfree(p.tail_pp);
  dfree(p.tail_pq);
  dfree(p.fold_rq);
  dfree(p.fold_qq);
  dfree(p.fold_rr);
  if (!p.tt_alias) { dfree(p.tt1); dfree(p.tt2); }
  p.tt1 = p.tt2 = nullptr;
  p.have_toep = false;
  if (p.graph_exec) (void)hipGraphExecDestroy(p.graph_exec);
  p.graph_exec = nullptr;
  p.graph_key.clear();
}

// the lean Toeplitz passes unless CGS_TOEP_LEAN=0
static bool toep_lean_enabled() {
  static const bool on = [] {
    const char* e = std::getenv("CGS_TOEP_LEAN");
    return !(e && std::atoi(e) == 0);
  }();
  return on;
}

// per-coil images X_c = crop(IFFT_2N(That .* FFT_2N(pad(S_c x)))) (Toeplitz embedding)
// crop and pad as one block per grid row over the row's 2..4 coil groups,
// loading group g + 1 while group g is transformed: the crop then sums all
// coils (one partial image), and the row kernels carry the folded CG tail.
// Measured at 256^2 x 32 coils before the fold: crop 13.65 -> 13.03 us, the row
// pad 13.58 -> 13.91 us against a block per group; CGS_TOEP_ROW=0 /
// CGS_TOEP_PAD_ROW=0 select a block per group of 16 coils.
static bool toep_row_shape(int C) { return C % 16 == 0 && C / 16 >= 2 && C / 16 <= 4; }
static bool toep_row_enabled(int C) {
  static const bool on = [] {
    const char* e = std::getenv("CGS_TOEP_ROW");
    return !(e && std::atoi(e) == 0);
  }();
  return on && toep_row_shape(C);
}
static bool toep_pad_row_enabled(int C) {   // CGS_TOEP_PAD_ROW=0: one block per coil group
  static const bool on = [] {
    const char* e = std::getenv("CGS_TOEP_PAD_ROW");
    return !(e && std::atoi(e) == 0);
  }();
  return on && toep_row_shape(C);
}

static int toeplitz_to_coils(Plan& p, const float2* x, const int* fl, const PUpd* pu, int* G_out,
                             bool* tail_out = nullptr, const cgs_solve_opts* fold = nullptr,
                             bool* folded = nullptr) {
  // inside the CG (pu set) the register FFT passes also produce the partials
  // of the fused CG tail: ||p||^2 per row (pad pass) and <p, q_partial> per
  // (coil group, row) (crop pass)
  int tail_parts = 0;
  const int d = p.d, C = p.C, F = p.F;
  const int64_t N0 = p.N[0], N1 = p.N[1], N2 = p.N[2];
  const int64_t M0 = 2 * N0, M1 = 2 * N1;
  // the last pass (axis 0, inner = C) crops and sums conj(S_c) X_c over the
  // block's coils into G partial images in p.Xc (store mode 2)
  auto pass = [&](int kind, const FftTab& tab, int sign, const float2* in, const float2* xin, float2* out, int nin,
                  int nout, int64_t inner, int64_t opf, int ld, const float* That, int64_t cells, double bytes,
                  int st = 0) {
    FftArgs a = fft_base(p, tab, sign);
    a.in = in; a.x = xin; a.out = out;
    a.nin = nin; a.nout = nout;
    a.inner = inner; a.outer_per_frame = opf; a.outers = opf * F;
    a.fdone = fl;
    if (ld == 4) apply_pupd(p, a, pu);
    if (That) { a.That = That; a.conv_cells = cells; a.pad_in = 1; a.crop_out = 1; }
    fft_finish(a);
    if (pu && tail_out && p.tail_pp && a.use_reg) {
      if (ld == 4) { a.pp_part = p.tail_pp; ++tail_parts; }
      if (st == 2) { a.pq_part = p.tail_pq; a.pdot = pu->pnew; ++tail_parts; }
    }
    if (st == 2) {
      const int G = (int)((inner + a.lb - 1) / a.lb);
      *G_out = G;
      bytes = (double)C * (nin * (double)(opf)) * 8 + (double)C * nout * (double)opf * 8 +
              (double)G * nout * (double)opf * 8;
    }
    Prof pr(p, kind, bytes * F);
    launch_fft(a, ld, st, p.stream);
  };
  // 3D on the lean kernels too when every pass has one: the axis-0 pad and
  // crop treat the N1 N2 rows of a frame like the 2D kernels' N1 rows, the
  // middle axis runs the half passes, the last axis the conv kernel
  const int64_t rows = (d == 2) ? N1 : N1 * N2;
  const bool lean3 = d == 3 && toep_lean_enabled() && toep_pad_crop_supported((int)N0, (int)rows, C) &&
                     toep_half_supported((int)N0, (int)N1, C) && toep_conv_supported((int)(M0 * N1), (int)N2, C);
  if (d == 2 || lean3) {
    // lean kernels (kernels_toep.hip) where the shape allows; pad and crop
    // switch together (both leave image positions unshifted along axis 0),
    // the conv pass is self-contained along its axis
    const bool lean_pc = lean3 || (toep_lean_enabled() && toep_pad_crop_supported((int)N0, (int)N1, C));
    const bool lean_cv = lean3 || (toep_lean_enabled() && toep_conv_supported((int)N0, (int)N1, C));
    ToepArgs ta{};
    ta.C = C; ta.N0 = (int)N0; ta.N1 = (int)rows; ta.F = F; ta.Ntot = p.Ntot;
    ta.S = p.S; ta.Sfs = p.Sfs; ta.fdone = fl;
    ta.row = 0;
    const bool want_tail = pu && tail_out && p.tail_pp;
    const double pad_bytes = ((double)(1 + C) * N0 * rows + (double)C * M0 * rows) * 8;
    if (lean_pc) {
      ToepArgs a = ta;
      a.row = toep_pad_row_enabled(C) ? 1 : 0;
      a.tw = p.ft[0].tw; a.x = x; a.out = p.tt1;
      if (pu) { a.r = pu->r; a.pold = pu->pold; a.pnew = pu->pnew; a.cgst = p.cgst; }
      if (want_tail) { a.pp_part = p.tail_pp; ++tail_parts; }
      if (want_tail && fold && folded) {
        // beta and the stop test of the previous iteration move into this pass
        a.rpart = cg_tail_rpartials(p);
        a.nrpart = cg_tail_blocks(p);
        a.tolkind = fold->tol_kind;
        a.tol = fold->tol;
        a.maxit = fold->maxit;
        a.cgw = p.cgst;
        a.hist = p.hist;
        a.hcap = p.hist_cap;
        a.fdone_w = p.fdone;
        *folded = true;
      }
      Prof pr(p, CGS_K_TOEP_PAD, pad_bytes * F);
      launch_toep_pad(a, p.stream);
    } else {
      pass(CGS_K_TOEP_PAD, p.ft[0], -1, nullptr, x, p.tt1, (int)N0, (int)M0, C, N1, 4, nullptr, 0, pad_bytes);
    }
    const double conv_bytes = 2.0 * C * M0 * N1 * 8 + 4.0 * M0 * M1;
    if (lean3) {
      // axis 1: tt1 (N1 rows per plane) -> tt2 (2 N1 rows), the conv along
      // axis 2 in tt2, axis 1 back into tt1
      ToepArgs h{};
      h.C = C; h.N0 = (int)N0; h.N1 = (int)N1; h.F = F; h.ppf = (int)N2; h.fdone = fl; h.tw = p.ft[1].tw;
      const double half_bytes = (double)C * M0 * (N1 + M1) * N2 * 8;
      {
        h.in = p.tt1; h.out = p.tt2;
        Prof pr(p, CGS_K_TOEP_PAD, half_bytes * F);
        launch_toep_half(h, +1, p.stream);
      }
      {
        ToepArgs a{};
        a.C = C; a.N0 = (int)(M0 * N1); a.N1 = (int)N2; a.F = F; a.Ntot = p.Ntot;
        a.fdone = fl; a.tw = p.ft[2].tw; a.out = p.tt2; a.That = p.That;
        Prof pr(p, CGS_K_TOEP_CONV, (2.0 * C * M0 * M1 * N2 * 8 + 4.0 * M0 * M1 * 2 * N2) * F);
        launch_toep_conv(a, p.stream);
      }
      {
        h.in = p.tt2; h.out = p.tt1;
        Prof pr(p, CGS_K_TOEP_CROP, half_bytes * F);
        launch_toep_half(h, -1, p.stream);
      }
    } else if (lean_cv) {
      ToepArgs a = ta;
      a.tw = p.ft[1].tw; a.out = p.tt1; a.That = p.That;
      Prof pr(p, CGS_K_TOEP_CONV, conv_bytes * F);
      launch_toep_conv(a, p.stream);
    } else {
      pass(CGS_K_TOEP_CONV, p.ft[1], -1, p.tt1, nullptr, p.tt1, (int)N1, (int)N1, C * M0, 1, 0, p.That, M0,
           conv_bytes);
    }
    if (lean_pc) {
      ToepArgs a = ta;
      a.tw = p.ft[0].tw; a.in = p.tt1; a.out = p.Xc;
      a.row = toep_row_enabled(C) ? 1 : 0;
      const int G = a.row ? 1 : C / 16;
      *G_out = G;
      if (want_tail) { a.pq_part = p.tail_pq; a.pdot = pu->pnew; ++tail_parts; }
      const double bytes = ((double)C * M0 * rows + (double)C * N0 * rows + (double)G * N0 * rows) * 8;
      Prof pr(p, CGS_K_TOEP_CROP, bytes * F);
      launch_toep_crop(a, p.stream);
    } else {
      pass(CGS_K_TOEP_CROP, p.ft[0], +1, p.tt1, nullptr, p.Xc, (int)M0, (int)N0, C, N1, 0, nullptr, 0, 0.0, 2);
    }
  } else {
    pass(CGS_K_TOEP_PAD, p.ft[0], -1, nullptr, x, p.tt1, (int)N0, (int)M0, C, N1 * N2, 4, nullptr, 0,
         ((double)(1 + C) * N0 * N1 * N2 + (double)C * M0 * N1 * N2) * 8);
    pass(CGS_K_TOEP_PAD, p.ft[1], -1, p.tt1, nullptr, p.tt2, (int)N1, (int)M1, C * M0, N2, 0, nullptr, 0,
         (double)C * M0 * (N1 + M1) * N2 * 8);
    // the conv along axis 2 on the lean kernel: its lines run over the
    // (2 N0 x 2 N1) cells x C coils like the 2D kernel's over 2 N0 x C, so it
    // takes N0' = 2 N0 N1 (2 N0' cells) and N1' = N2
    const double conv3_bytes = 2.0 * C * M0 * M1 * N2 * 8 + 4.0 * M0 * M1 * 2 * N2;
    if (toep_lean_enabled() && toep_conv_supported((int)(M0 * N1), (int)N2, C)) {
      ToepArgs a{};
      a.C = C; a.N0 = (int)(M0 * N1); a.N1 = (int)N2; a.F = F; a.Ntot = p.Ntot;
      a.fdone = fl; a.tw = p.ft[2].tw; a.out = p.tt2; a.That = p.That;
      Prof pr(p, CGS_K_TOEP_CONV, conv3_bytes * F);
      launch_toep_conv(a, p.stream);
    } else {
      pass(CGS_K_TOEP_CONV, p.ft[2], -1, p.tt2, nullptr, p.tt2, (int)N2, (int)N2, C * M0 * M1, 1, 0, p.That, M0 * M1,
           conv3_bytes);
    }
    pass(CGS_K_TOEP_CROP, p.ft[1], +1, p.tt2, nullptr, p.tt1, (int)M1, (int)N1, C * M0, N2, 0, nullptr, 0,
         (double)C * M0 * (M1 + N1) * N2 * 8);
    pass(CGS_K_TOEP_CROP, p.ft[0], +1, p.tt1, nullptr, p.Xc, (int)M0, (int)N0, C, N1 * N2, 0, nullptr, 0, 0.0, 2);
  }
  if (tail_out) *tail_out = (tail_parts == 2);
  return CGS_OK;
}

// q = E^H W E x + lambda x with the Toeplitz embedding (and partial <x, q>)
static int op_normal_toeplitz(Plan& p, const float2* x, float2* qo, float lambda, double* partials,
                              const int* fl) {
  int G = 0;
  CHK(toeplitz_to_coils(p, x, fl, nullptr, &G));
  Prof pr(p, CGS_K_COIL, ((double)G + 3.0) * p.Ntot * p.F * 8);
  launch_partial_combine(p, p.Xc, G, qo, x, lambda, partials, fl, p.stream);
  return CGS_OK;
}

// one application of the CG operator (E^H W E + lambda) for p, with <p, q> partials
static int op_normal(Plan& p, int normal_op, const float2* x, float2* qo, float lambda, double* partials,
                     const int* fl) {
  if (normal_op == CGS_NORMAL_TOEPLITZ) return op_normal_toeplitz(p, x, qo, lambda, partials, fl);
  CHK(op_E(p, x, p.ybuf[1], fl));
  return op_EH_true(p, p.ybuf[1], qo, x, lambda, partials, fl);
}

// The folded CG iteration (2D, the row pad/crop kernels): three passes, no
// tail kernel.  The pad applies the previous iteration's tail to its row (alpha
// from the crop's <p, q>, x += alpha p, r -= alpha q) with ||r_new||^2 from the
// recurrence rn - 2 alpha <r, q> + alpha^2 |q|^2, forms beta and the stop test,
// then p = r + beta p and the first FFT pass; the crop writes q = A p + lambda p
// with its per-row partials and passes the pad's stop flag on.
static bool tailfold_ok(const Plan& p, const cgs_solve_opts& o) {
  const int C = p.C;
  return p.tailfold && p.d == 2 && o.normal_op == CGS_NORMAL_TOEPLITZ && o.eh_mode == CGS_EH_TRUE &&
         toep_lean_enabled() && toep_pad_crop_supported((int)p.N[0], (int)p.N[1], C) &&
         toep_conv_supported((int)p.N[0], (int)p.N[1], C) && toep_pad_row_enabled(C) && toep_row_enabled(C) &&
         p.fold_rq && p.fold_qq && p.fold_rr;
}

static int cg_iteration_tailfold(Plan& p, float2* x, float2* r, const float2* pold, float2* pnew, float2* q,
                                 const cgs_solve_opts& o) {
  const int C = p.C, F = p.F;
  const int64_t N0 = p.N[0], N1 = p.N[1], M0 = 2 * N0;
  ToepArgs ta{};
  ta.C = C; ta.N0 = (int)N0; ta.N1 = (int)N1; ta.F = F; ta.Ntot = p.Ntot;
  ta.S = p.S; ta.Sfs = p.Sfs; ta.fdone = p.fdone;
  {
    ToepArgs a = ta;
    a.row = 1;
    a.tw = p.ft[0].tw; a.out = p.tt1;
    a.r = r; a.pold = pold; a.pnew = pnew; a.cgst = p.cgst;
    a.tailfold = 1; a.xs = x; a.rs = r; a.qv = q;
    a.fpq = p.tail_pq; a.frq = p.fold_rq; a.fqq = p.fold_qq; a.frr = p.fold_rr; a.nrpart = (int)N1;
    a.tolkind = o.tol_kind; a.tol = o.tol; a.maxit = o.maxit;
    a.cgw = p.cgst; a.hist = p.hist; a.hcap = p.hist_cap;
    // reads r, p_old, q, x and the maps, writes x, r, p_new and the grid rows
    const double bytes = ((double)(7 + C) * N0 * N1 + (double)C * M0 * N1) * 8;
    Prof pr(p, CGS_K_TOEP_PAD, bytes * F);
    launch_toep_pad(a, p.stream);
  }
  {
    ToepArgs a = ta;
    a.tw = p.ft[1].tw; a.out = p.tt1; a.That = p.That;
    Prof pr(p, CGS_K_TOEP_CONV, (2.0 * C * M0 * N1 * 8 + 4.0 * M0 * 2 * N1) * F);
    launch_toep_conv(a, p.stream);
  }
  {
    ToepArgs a = ta;
    a.row = 1;
    a.tw = p.ft[0].tw; a.in = p.tt1; a.out = q;
    a.pq_part = p.tail_pq; a.pdot = pnew;
    a.tailfold = 1; a.lam = (float)o.lambda; a.rdot = r; a.rq_part = p.fold_rq; a.qq_part = p.fold_qq;
    a.rr_part = p.fold_rr;
    a.cgst = p.cgst; a.cgw = p.cgst; a.fdone_w = p.fdone;
    const double bytes = ((double)C * M0 * N1 + (double)C * N0 * N1 + 3.0 * N0 * N1) * 8;
    Prof pr(p, CGS_K_TOEP_CROP, bytes * F);
    launch_toep_crop(a, p.stream);
  }
  p.beta_folded = false;
  p.tailfold_used = true;
  return CGS_OK;
}

// One fused CG iteration of the true-adjoint solve: p_new = r + beta p_old
// inside the first FFT pass, q = (E^H W E + lambda) p_new with alpha formed by
// the coil combine's last block, then x, r, beta and the stop test.
static int cg_iteration_fused(Plan& p, int normal_op, float2* x, float2* r, const float2* pold, float2* pnew,
                              float2* q, const cgs_solve_opts& o) {
  if (tailfold_ok(p, o)) return cg_iteration_tailfold(p, x, r, pold, pnew, q, o);
  PUpd pu;
  pu.r = r;
  pu.pold = pold;
  pu.pnew = pnew;
  if (normal_op == CGS_NORMAL_TOEPLITZ) {
    int G = 0;
    bool tail = false, folded = false;
    CHK(toeplitz_to_coils(p, nullptr, p.fdone, &pu, &G, &tail, &o, &folded));
    if (tail) {
      // alpha from the FFT passes' partials, q formed in place: no combine
      // kernel; with the lean pad pass beta is formed there (next iteration)
      const int opf = (int)(p.Ntot / p.N[0]);
      Prof pr(p, CGS_K_CG, ((double)G + 6.0) * 8.0 * p.F * p.Ntot);
      launch_cg_tail(p, x, r, pnew, p.Xc, G, (float)o.lambda, p.tail_pq, G * opf, p.tail_pp, opf, folded,
                     p.stream);
      if (!folded) launch_cg_beta_after_xr(p, cg_tail_blocks(p), o.tol_kind, o.tol, o.maxit, p.stream);
      p.beta_folded = folded;
      return CGS_OK;
    }
    Prof pr(p, CGS_K_COIL, ((double)G + 3.0) * p.Ntot * p.F * 8);
    launch_partial_combine(p, p.Xc, G, q, pnew, (float)o.lambda, p.partials, p.fdone, p.stream);
  } else {
    CHK(op_E(p, nullptr, p.ybuf[1], p.fdone, &pu));
    CHK(gridding(p, p.ybuf[1], true, p.fdone));
    CHK(fft_adjoint(p, p.Xc, p.fdone));
    Prof pr(p, CGS_K_COIL, coil_bytes(p, true));
    launch_coil_combine(p, p.Xc, q, pnew, (float)o.lambda, false, p.partials, p.fdone, p.stream);
  }
  {
    Prof pr(p, CGS_K_CG, 5.0 * 8.0 * p.F * p.Ntot);
    launch_cg_xr_alpha(p, x, r, pnew, q, p.stream);
    launch_cg_beta_after_xr(p, kRedBlocks, o.tol_kind, o.tol, o.maxit, p.stream);
  }
  return CGS_OK;
}

// One CG iteration on separate kernels in codeare's order (CGLS.hpp:48-56):
// q = (A + lambda) p with <p, q> partials, alpha, x and r with ||r||^2
// partials, beta and the stop test, p = r + beta p.  The reference E^H takes
// this path, and so does a joint solve (scalars summed over all frames).
static int cg_iteration_plain(Plan& p, float2* x, float2* r, float2* pv, float2* q, const cgs_solve_opts& o,
                              bool joint) {
  if (o.eh_mode == CGS_EH_TRUE) {
    CHK(op_normal(p, o.normal_op, pv, q, (float)o.lambda, p.partials, p.fdone));
  } else {
    CHK(op_E(p, pv, p.ybuf[1], p.fdone));
    CHK(op_EH(p, o.eh_mode, p.ybuf[1], q, o.ftiter, o.fteps, pv, (float)o.lambda, p.partials, p.fdone));
  }
  Prof pr(p, CGS_K_CG, 7.0 * 8.0 * p.F * p.Ntot);
  launch_cg_alpha(p, p.partials, joint, p.stream);
  launch_cg_xr(p, x, r, pv, q, p.partials, p.stream);
  launch_cg_beta(p, p.partials, o.tol_kind, o.tol, o.maxit, joint, p.stream);
  launch_cg_p(p, pv, r, p.stream);
  return CGS_OK;
}

// ---- CG solve -------------------------------------------------------------------------
static double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

static int solve(Plan& p, const float2* y_ext, const float2* x0, float2* xout, const cgs_solve_opts& o,
                 cgs_solve_report* rep, std::chrono::steady_clock::time_point t_start) {
  const int F = p.F;
  const bool joint = o.joint != 0 && F > 1;
  p.beta_folded = false;
  p.tailfold_used = false;
  std::vector<double>& done_ms = p.done_ms;
  std::fill(done_ms.begin(), done_ms.end(), -1.0);
  const size_t ib = sizeof(float2) * (size_t)F * p.Ntot;
  float2 *x = p.img[0], *r = p.img[1], *pv = p.img[2], *q = p.img[3];
  double* partA = p.partials;
  double* partB = p.partials + 512 * F;
  if (o.maxit + 1 > p.hist_cap) { set_error("residual history smaller than maxit + 1"); return CGS_ERR_STATE; }
  p.last_hist = p.hist;
  p.last_hist_cap = p.hist_cap;
  // fills are kernels (no memset nodes where a caller captures the stream)
  launch_zero(p.hist, sizeof(double) * F * p.hist_cap, p.stream);
  launch_fill_int(p.fdone, 0, F, p.stream);
  {
    Prof pr(p, CGS_K_OTHER, 2.0 * sizeof(float2) * p.Mtot * p.C);
    launch_y_ext2int(p, y_ext, p.ybuf[0], p.C, -1, p.stream);
  }
  // b = E^H y  -> r
  CHK(op_EH(p, o.eh_mode, p.ybuf[0], r, o.ftiter, o.fteps, nullptr, 0.f, nullptr, p.fdone));
  const bool warm = x0 != nullptr;
  if (o.eh_mode == CGS_EH_TRUE && o.normal_op == CGS_NORMAL_TOEPLITZ) CHK(toeplitz_prepare(p));
  if (warm) {
    HIPCHK(hipMemcpyAsync(x, x0, ib, hipMemcpyDeviceToDevice, p.stream));
    if (o.eh_mode == CGS_EH_TRUE) {
      CHK(op_normal(p, o.normal_op, x, q, (float)o.lambda, nullptr, p.fdone));
    } else {
      CHK(op_E(p, x, p.ybuf[1], p.fdone));
      CHK(op_EH(p, o.eh_mode, p.ybuf[1], q, o.ftiter, o.fteps, x, (float)o.lambda, nullptr, p.fdone));
    }
    launch_sqnorm_partials(p, r, partA, p.stream);             // ||b||^2
    launch_axpby_img(p, r, r, 1.f, q, -1.f, p.stream);          // r = b - (A + lambda) x0
    launch_sqnorm_partials(p, r, partB, p.stream);
  } else {
    launch_zero(x, ib, p.stream);
    launch_sqnorm_partials(p, r, partA, p.stream);
  }
  launch_cg_init(p, partA, partB, o.tol_kind, o.tol, o.maxit, warm, joint, p.stream);
  HIPCHK(hipMemcpyAsync(pv, r, ib, hipMemcpyDeviceToDevice, p.stream));
  CHK(check_launch());

  hipEvent_t* ev = p.poll_ev;   // made at plan creation
  int status = CGS_OK;
  float2* pbuf[2] = {pv, p.img[6]};   // p double buffer for the fused direction update
  int step = 1;
  if (o.graph_iters > 0 && o.eh_mode == CGS_EH_TRUE) {
    // capture an even number of fused iterations once and replay the graph
    step = o.graph_iters + (o.graph_iters & 1);
    char key[256];
    snprintf(key, sizeof(key), "%d|%d|%.17g|%d|%.17g|%d|%p|%p|%d|%p|%d|%d", step, o.normal_op, o.lambda, o.tol_kind,
             o.tol, o.maxit, (void*)p.That, (void*)p.stream, joint ? 1 : 0, (void*)p.hist, p.hist_cap,
             tailfold_ok(p, o) ? 1 : 0);
    if (p.graph_exec && p.graph_key == key) {
      p.beta_folded = p.graph_folded;
      p.tailfold_used = p.graph_tailfold;
    }
    if (!p.graph_exec || p.graph_key != key) {
      if (p.graph_exec) (void)hipGraphExecDestroy(p.graph_exec);
      p.graph_exec = nullptr;
      hipGraph_t g = nullptr;
      p.capturing = true;
      HIPCHK(hipStreamBeginCapture(p.stream, hipStreamCaptureModeThreadLocal));
      int rc = CGS_OK;
      for (int k = 0; k < step && rc == CGS_OK; ++k)
        rc = joint ? cg_iteration_plain(p, x, r, pv, q, o, true)
                   : cg_iteration_fused(p, o.normal_op, x, r, pbuf[k & 1], pbuf[(k + 1) & 1], q, o);
      hipError_t ec = hipStreamEndCapture(p.stream, &g);
      p.capturing = false;
      if (rc != CGS_OK) { if (g) (void)hipGraphDestroy(g); return rc; }
      if (ec != hipSuccess) { set_error(std::string("graph capture failed: ") + hipGetErrorString(ec)); return CGS_ERR_HIP; }
      hipError_t ei = hipGraphInstantiate(&p.graph_exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ei != hipSuccess) { p.graph_exec = nullptr; set_error(std::string("graph instantiate failed: ") + hipGetErrorString(ei)); return CGS_ERR_HIP; }
      p.graph_key = key;
      p.graph_folded = p.beta_folded;
      p.graph_tailfold = p.tailfold_used;
    }
  }
  int launched = 0;   // iterations issued (p_new of the last one is pbuf[launched & 1])
  for (int it = 0; it < o.maxit; it += step) {
    launched += step;
    if (step > 1) {
      HIPCHK(hipGraphLaunch(p.graph_exec, p.stream));
    } else if (o.eh_mode == CGS_EH_TRUE && !joint) {
      CHK(cg_iteration_fused(p, o.normal_op, x, r, pbuf[it & 1], pbuf[(it + 1) & 1], q, o));
    } else {
      CHK(cg_iteration_plain(p, x, r, pv, q, o, joint));
    }
    CHK(check_launch());
    // pipelined convergence test: read the flags of the previous iteration
    const int par = (it / step) & 1;
    HIPCHK(hipMemcpyAsync(p.h_done + par * F, p.fdone, sizeof(int) * F, hipMemcpyDeviceToHost, p.stream));
    HIPCHK(hipEventRecord(ev[par], p.stream));
    if (it >= step) {
      HIPCHK(hipEventSynchronize(ev[par ^ 1]));
      bool all = true;
      const double now = ms_since(t_start);
      for (int f = 0; f < F; ++f) {
        const bool d = p.h_done[(par ^ 1) * F + f] != 0;
        if (d && done_ms[f] < 0) done_ms[f] = now;
        all &= d;
      }
      if (all) break;
    }
  }
  // with beta folded into the pad pass the last iteration's stop test is
  // still pending when the loop ends at maxit: commit it (frames already done
  // are skipped)
  if (p.beta_folded && o.eh_mode == CGS_EH_TRUE && o.normal_op == CGS_NORMAL_TOEPLITZ)
    launch_cg_beta_after_xr(p, cg_tail_blocks(p), o.tol_kind, o.tol, o.maxit, p.stream);
  // with the whole tail folded, the last iteration's x, r update and stop test
  // are pending for every frame still running
  if (p.tailfold_used && !joint)
    launch_cg_fold_final(p, x, r, pbuf[launched & 1], q, p.tail_pq, p.fold_rq, p.fold_qq, p.fold_rr, (int)p.N[1],
                         o.tol_kind, o.tol, o.maxit, p.stream);
  HIPCHK(hipMemcpyAsync(xout, x, ib, hipMemcpyDeviceToDevice, p.stream));
  HIPCHK(hipMemcpyAsync(p.h_cgst, p.cgst, sizeof(CgState) * F, hipMemcpyDeviceToHost, p.stream));
  launch_fill_int(p.fdone, 0, F, p.stream);
  HIPCHK(hipStreamSynchronize(p.stream));
  const double t_end = ms_since(t_start);
  for (int f = 0; f < F; ++f)
    if (done_ms[f] < 0) done_ms[f] = t_end;
  int maxi = 0;
  for (int f = 0; f < F; ++f) {
    maxi = std::max(maxi, p.h_cgst[f].it);
    if (p.h_cgst[f].nan) status = CGS_ERR_NAN;
  }
  if (rep) {
    rep->nframes = F;
    rep->max_iters = maxi;
    rep->status = status;
    for (int f = 0; f < F && f < 1024; ++f) {
      rep->iters[f] = p.h_cgst[f].it;
      rep->final_rel[f] = p.h_cgst[f].rel;
      rep->done_ms[f] = done_ms[f];
    }
  }
  return CGS_OK;
}

}  // namespace cgs

// =========================================================================================
// C ABI
// =========================================================================================
using namespace cgs;

#define GUARD_BEGIN try {
#define GUARD_END                                     \
  }                                                   \
  catch (const std::exception& e) {                   \
    set_error(std::string("exception: ") + e.what()); \
    return CGS_ERR_ARG;                               \
  }                                                   \
  catch (...) {                                       \
    set_error("unknown exception");                   \
    return CGS_ERR_ARG;                               \
  }

extern "C" {

const char* cgs_status_string(int s) {
  switch (s) {
    case CGS_OK: return "ok";
    case CGS_ERR_ARG: return "bad argument";
    case CGS_ERR_STATE: return "call out of order";
    case CGS_ERR_HIP: return "HIP runtime error";
    case CGS_ERR_ALLOC: return "device allocation failed";
    case CGS_ERR_UNSUPPORTED: return "unsupported parameters";
    case CGS_ERR_NAN: return "NaN in solver scalars";
    default: return "unknown status";
  }
}

const char* cgs_last_error(void) { return g_err.c_str(); }

int cgs_plan_create(cgs_plan* plan, int d, const int* N, int64_t M_total, double sigma, int m, int ncoils,
                    int nframes, const int64_t* frame_offsets, unsigned flags) {
  GUARD_BEGIN
  if (!plan) { set_error("null plan pointer"); return CGS_ERR_ARG; }
  *plan = nullptr;
  cgs_plan h = new cgs_plan_s();
  int rc = create_plan(h->p, d, N, M_total, sigma, m, ncoils, nframes, frame_offsets, flags);
  if (rc != CGS_OK) {
    destroy_plan(h->p);
    delete h;
    return rc;
  }
  *plan = h;
  return CGS_OK;
  GUARD_END
}

int cgs_plan_destroy(cgs_plan plan) {
  GUARD_BEGIN
  if (!plan) return CGS_OK;
  destroy_plan(plan->p);
  delete plan;
  return CGS_OK;
  GUARD_END
}

int cgs_set_stream(cgs_plan plan, void* stream) {
  GUARD_BEGIN
  if (!plan) return CGS_ERR_ARG;
  Plan& p = plan->p;
  HIPCHK(hipStreamSynchronize(p.stream));
  if (p.own_stream) (void)hipStreamDestroy(p.stream);
  p.stream = (hipStream_t)stream;
  p.own_stream = false;
  if (p.b0q) p.b0q->stream = p.stream;
  return CGS_OK;
  GUARD_END
}

int cgs_sync(cgs_plan plan) {
  GUARD_BEGIN
  if (!plan) return CGS_ERR_ARG;
  HIPCHK(hipStreamSynchronize(plan->p.stream));
  return CGS_OK;
  GUARD_END
}

int cgs_set_trajectory(cgs_plan plan, const float* k) {
  GUARD_BEGIN
  if (!plan || !k) { set_error("null argument"); return CGS_ERR_ARG; }
  Plan& p = plan->p;
  HIPCHK(hipStreamSynchronize(p.stream));
  CHK(build_samples(p, k));
  p.h_k.assign(k, k + (size_t)p.d * p.Mtot);
  toep_drop(p);
  if (!p.h_b0.empty()) CHK(b0_build(p));   // the off-resonance helper follows the trajectory
  p.have_k = true;
  p.have_w = false;   // weights follow the sample order (NFFT.hpp:258-296: Weights after KSpace)
  if (p.flags & CGS_AUTOTUNE) CHK(autotune(p, 5, true, nullptr));
  return CGS_OK;
  GUARD_END
}

// ---- launch-shape autotuner ------------------------------------------------------------
namespace cgs {
struct TuneChoice {
  int vpi, cgr, icg;
  double grid_ms, interp_ms, def_grid_ms, def_interp_ms;
  bool mfma;
  bool sweep;
  bool bin;
  int ilanes;
};
static std::mutex g_tune_mu;
static std::vector<std::pair<std::string, TuneChoice>> g_tune_cache;

static std::string tune_key(const Plan& p) {
  int64_t per = std::max<int64_t>(1, p.Mtot / p.F), b = 1;
  while (b < per) b <<= 1;
  char key[160];
  snprintf(key, sizeof(key), "%d|%d|%d|%d|%d|%d|%lld|%d", p.d, p.n[0], p.n[1], p.n[2], p.C, p.F, (long long)b, p.m);
  return key;
}

static int time_kernel(Plan& p, int reps, bool grid, double* ms) {
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  // one untimed run, then `reps` timed ones
  for (int r = 0; r <= reps; ++r) {
    if (r == 1) HIPCHK(hipEventRecord(e0, p.stream));
    if (grid) CHK(gridding(p, p.ybuf[0], false, nullptr));
    else CHK(interp(p, p.ybuf[1], nullptr));
  }
  HIPCHK(hipEventRecord(e1, p.stream));
  HIPCHK(hipEventSynchronize(e1));
  float t = 0.f;
  HIPCHK(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  CHK(check_launch());
  *ms = t / std::max(1, reps);
  return CGS_OK;
}

static int apply_vpi(Plan& p, int vpi) {
  if (p.visits_per_item == vpi) return CGS_OK;
  p.visits_per_item = vpi;
  HIPCHK(hipStreamSynchronize(p.stream));
  CHK(build_samples(p, p.h_k.data(), false));
  if (p.have_w) {   // weights are stored in the sorted sample order
    std::vector<float> ws((size_t)p.Mtot);
    for (int64_t s = 0; s < p.Mtot; ++s) ws[s] = p.h_w[p.h_perm[s]];
    HIPCHK(hipMemcpy(p.w, ws.data(), sizeof(float) * ws.size(), hipMemcpyHostToDevice));
  }
  return CGS_OK;
}

static int autotune(Plan& p, int reps, bool use_cache, cgs_tune_report* rep) {
  if (!p.have_k || p.h_k.empty()) { set_error("autotune needs the trajectory"); return CGS_ERR_STATE; }
  const std::string key = tune_key(p);
  TuneChoice best{};
  bool cached = false;
  if (use_cache) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (auto& kv : g_tune_cache)
      if (kv.first == key) { best = kv.second; cached = true; }
  }
  if (!cached) {
    reps = std::max(1, std::min(reps, 50));
    HIPCHK(hipMemsetAsync(p.ybuf[0], 0, sizeof(float2) * (size_t)p.Mtot * p.C, p.stream));
    HIPCHK(hipMemsetAsync(p.grid, 0, sizeof(float2) * (size_t)p.ntot * p.C * p.F, p.stream));
    const bool tiles8 = p.T == 8;
    const bool reg = p.d == 2 && tiles8 && p.use_grid_reg;
    std::vector<int> vpis = (p.d == 2) ? std::vector<int>{256, 512, 1024} : std::vector<int>{1024, 4096, 16384};
    std::vector<int> cgrs = (reg && p.C > 8) ? std::vector<int>{8, 16} : std::vector<int>{p.grid_cgr};
    std::vector<int> icgs = (p.C > 8) ? std::vector<int>{8, 16} : std::vector<int>{p.interp_cg};
    const int vpi0 = p.visits_per_item, cgr0 = p.grid_cgr, icg0 = p.interp_cg, il0 = p.interp_lanes;
    const bool mfma0 = p.use_grid_mfma;
    const bool sweep0 = p.use_grid_sweep && p.sw_items != nullptr;
    const bool bin0 = p.use_grid_bin && p.bin_ids != nullptr;
    double g0 = 0, i0 = 0;
    CHK(time_kernel(p, reps, true, &g0));
    CHK(time_kernel(p, reps, false, &i0));
    best = TuneChoice{vpi0, cgr0, icg0, g0, i0, g0, i0, mfma0, sweep0, bin0, il0};
    p.use_grid_bin = false;
    if (p.sw_items) {   // the line sweep (independent of the visit-list chunking)
      p.use_grid_sweep = true;
      double t = 0;
      CHK(time_kernel(p, reps, true, &t));
      if (t < best.grid_ms) { best.grid_ms = t; best.sweep = true; best.bin = false; }
    }
    p.use_grid_sweep = false;
    if (p.bin_ids) {    // the binned scatter (likewise)
      p.use_grid_bin = true;
      double t = 0;
      CHK(time_kernel(p, reps, true, &t));
      if (t < best.grid_ms) { best.grid_ms = t; best.bin = true; best.sweep = false; }
    }
    p.use_grid_bin = false;   // the tile kernels below
    std::vector<int> kinds = tiles8 ? std::vector<int>{1, 0} : std::vector<int>{0};   // 1: matrix cores
    for (int vpi : vpis) {
      CHK(apply_vpi(p, vpi));
      for (int kind : kinds) {
        p.use_grid_mfma = kind == 1;
        for (int cgr : (kind == 1 ? std::vector<int>{p.grid_cgr} : cgrs)) {
          p.grid_cgr = cgr;
          double t = 0;
          CHK(time_kernel(p, reps, true, &t));
          if (t < best.grid_ms) {
            best.grid_ms = t; best.vpi = vpi; best.cgr = cgr; best.mfma = kind == 1; best.sweep = false;
            best.bin = false;
          }
        }
      }
    }
    // the gather does not depend on the visit lists; the LDS gathers try both
    // lane layouts
    const bool cs3 = p.d == 3 && p.use_interp_lds && p.taps == 8 && p.T == 8;
    const std::vector<int> lanes = (p.d == 2 && p.use_interp_lds) ? std::vector<int>{4, 3, 1, 0}
                                   : cs3 ? std::vector<int>{2, 0} : std::vector<int>{il0};
    for (int il : lanes)
      for (int icg : icgs) {
        p.interp_cg = icg;
        p.interp_lanes = il;
        double t = 0;
        CHK(time_kernel(p, reps, false, &t));
        if (t < best.interp_ms) { best.interp_ms = t; best.icg = icg; best.ilanes = il; }
      }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache.emplace_back(key, best);
  }
  CHK(apply_vpi(p, best.vpi));
  p.grid_cgr = best.cgr;
  p.interp_cg = best.icg;
  p.interp_lanes = best.ilanes;
  p.use_grid_mfma = best.mfma;
  p.use_grid_sweep = best.sweep && p.sw_items != nullptr;
  p.use_grid_bin = best.bin && p.bin_ids != nullptr;
  // captured graphs hold the old launch shapes
  if (p.graph_exec) (void)hipGraphExecDestroy(p.graph_exec);
  p.graph_exec = nullptr;
  p.graph_key.clear();
  if (rep) {
    rep->visits_per_item = best.vpi;
    rep->grid_coils = best.cgr;
    rep->interp_coils = best.icg;
    rep->from_cache = cached ? 1 : 0;
    rep->grid_ms = best.grid_ms;
    rep->interp_ms = best.interp_ms;
    rep->default_grid_ms = best.def_grid_ms;
    rep->default_interp_ms = best.def_interp_ms;
    rep->grid_mfma = best.mfma ? 1 : 0;
    rep->grid_sweep = best.sweep ? 1 : 0;
    rep->grid_bin = best.bin ? 1 : 0;
    rep->interp_lanes = best.ilanes;
  }
  return CGS_OK;
}
}  // namespace cgs

int cgs_autotune(cgs_plan plan, int reps, int use_cache, cgs_tune_report* report) {
  GUARD_BEGIN
  if (!plan) { set_error("null argument"); return CGS_ERR_ARG; }
  return autotune(plan->p, reps, use_cache != 0, report);
  GUARD_END
}

int cgs_set_b0(cgs_plan plan, const float* b0, const float* t) {
  GUARD_BEGIN
  if (!plan) { set_error("null argument"); return CGS_ERR_ARG; }
  Plan& p = plan->p;
  HIPCHK(hipStreamSynchronize(p.stream));
  if (!b0) {   // back to the plain NFFT
    b0_free(p);
    p.h_b0.clear();
    p.h_t.clear();
    return CGS_OK;
  }
  if (!t) { set_error("timing required with b0"); return CGS_ERR_ARG; }
  if (p.d != 2 || p.F != 1) { set_error("the off-resonance model is 2D, one frame (codeare's mri_inh_3d plan)"); return CGS_ERR_UNSUPPORTED; }
  if (!p.have_k) { set_error("set the trajectory before b0 and timing"); return CGS_ERR_STATE; }
  for (int64_t i = 0; i < p.Ntot; ++i)
    if (!std::isfinite(b0[i])) { set_error("b0 is not finite"); return CGS_ERR_ARG; }
  for (int64_t j = 0; j < p.Mtot; ++j)
    if (!std::isfinite(t[j])) { set_error("timing is not finite"); return CGS_ERR_ARG; }
  p.h_b0.assign(b0, b0 + (size_t)p.Ntot);
  p.h_t.assign(t, t + (size_t)p.Mtot);
  const int rc = b0_build(p);
  if (rc != CGS_OK) { b0_free(p); p.h_b0.clear(); p.h_t.clear(); }
  return rc;
  GUARD_END
}

int cgs_b0_info(cgs_plan plan, int* N3, double* omega, double* ts) {
  GUARD_BEGIN
  if (!plan) { set_error("null argument"); return CGS_ERR_ARG; }
  const Plan& p = plan->p;
  if (!p.b0q) { set_error("no off-resonance model on this plan"); return CGS_ERR_STATE; }
  if (N3) *N3 = p.b0_N3;
  if (omega) *omega = p.b0_omega;
  if (ts) *ts = p.b0_ts;
  return CGS_OK;
  GUARD_END
}

int cgs_set_weights(cgs_plan plan, const float* w) {
  GUARD_BEGIN
  if (!plan || !w) { set_error("null argument"); return CGS_ERR_ARG; }
  Plan& p = plan->p;
  if (!p.have_k) { set_error("set the trajectory before the weights (NFFT::Weights after KSpace)"); return CGS_ERR_STATE; }
  std::vector<float> ws((size_t)p.Mtot);
  for (int64_t s = 0; s < p.Mtot; ++s) ws[s] = w[p.h_perm[s]];
  HIPCHK(hipStreamSynchronize(p.stream));
  HIPCHK(hipMemcpy(p.w, ws.data(), sizeof(float) * ws.size(), hipMemcpyHostToDevice));
  CHK(sweep_weights(p, ws));
  p.h_w.assign(w, w + (size_t)p.Mtot);
  CHK(b0_weights(p));
  toep_drop(p);
  p.have_w = true;
  return CGS_OK;
  GUARD_END
}

int cgs_set_sensitivities(cgs_plan plan, const void* S) {
  GUARD_BEGIN
  if (!plan || !S) { set_error("null argument"); return CGS_ERR_ARG; }
  Plan& p = plan->p;
  const float2* s = (const float2*)S;
  // host (N^d, nm, C) column-major -> device [frame][pixel][coil]
  const int64_t nm = p.Sfs ? p.F : 1;
  std::vector<float2> si((size_t)p.Ntot * p.C * nm);
  std::vector<float> I((size_t)p.Ntot * nm);
  for (int64_t z = 0; z < nm; ++z)
    for (int64_t i = 0; i < p.Ntot; ++i) {
      float acc = 0.f;   // IntensityMap accumulates in the maps' precision (MRI.hpp:21-26)
      for (int c = 0; c < p.C; ++c) {
        const float2 v = s[i + p.Ntot * (z + nm * c)];
        si[(z * p.Ntot + i) * p.C + c] = v;
        acc += v.x * v.x + v.y * v.y;
      }
      I[z * p.Ntot + i] = (float)(1.0 / (std::sqrt((double)acc) + 1.e-9));
    }
  HIPCHK(hipStreamSynchronize(p.stream));
  HIPCHK(hipMemcpy(p.S, si.data(), sizeof(float2) * si.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(p.Imap, I.data(), sizeof(float) * I.size(), hipMemcpyHostToDevice));
  p.have_S = true;
  return CGS_OK;
  GUARD_END
}

int cgs_set_sensitivities_device(cgs_plan plan, const void* S) {
  GUARD_BEGIN
  if (!plan || !S) { set_error("null argument"); return CGS_ERR_ARG; }
  Plan& p = plan->p;
  const int64_t nm = p.Sfs ? p.F : 1;
  launch_maps_load((const float2*)S, p.S, p.Imap, p.Ntot, nm, p.C, p.stream);
  CHK(check_launch());
  p.have_S = true;
  return CGS_OK;
  GUARD_END
}

int cgs_E(cgs_plan plan, const void* x, void* y) {
  GUARD_BEGIN
  if (!plan || !x || !y) { set_error("null argument"); return CGS_ERR_ARG; }
  Plan& p = plan->p;
  CHK(require_ready(p, true));
  const float2* xd;
  CHK(stage_in(p, x, (float2*)p.ext_b, sizeof(float2) * p.Ntot * p.F, &xd));
  float2* yd = (p.flags & CGS_DEVICE_IO) ? (float2*)y : (float2*)p.ext_a;
  CHK(op_E(p, xd, p.ybuf[1], nullptr));
  {
    Prof pr(p, CGS_K_OTHER, 2.0 * sizeof(float2) * p.Mtot * p.C);
    launch_y_int2ext(p, p.ybuf[1], yd, p.C, -1, p.stream);
  }
  CHK(check_launch());
  CHK(stage_out(p, y, yd, sizeof(float2) * p.Mtot * p.C));
  return CGS_OK;
  GUARD_END
}

int cgs_EH(cgs_plan plan, const void* y, void* x, int mode, int ftiter, double fteps) {
  GUARD_BEGIN
  if (!plan || !y || !x) { set_error("null argument"); return CGS_ERR_ARG; }
  if (mode != CGS_EH_TRUE && mode != CGS_EH_REF) { set_error("bad E^H mode"); return CGS_ERR_ARG; }
  if (ftiter < 0) { set_error("ftiter must be >= 0"); return CGS_ERR_ARG; }
  Plan& p = plan->p;
  CHK(require_ready(p, true));
  const float2* yd;
  CHK(stage_in(p, y, (float2*)p.ext_a, sizeof(float2) * p.Mtot * p.C, &yd));
  float2* xd = (p.flags & CGS_DEVICE_IO) ? (float2*)x : (float2*)p.ext_b;
  {
    Prof pr(p, CGS_K_OTHER, 2.0 * sizeof(float2) * p.Mtot * p.C);
    launch_y_ext2int(p, yd, p.ybuf[0], p.C, -1, p.stream);
  }
  CHK(op_EH(p, mode, p.ybuf[0], xd, ftiter, fteps, nullptr, 0.f, nullptr, nullptr));
  CHK(check_launch());
  CHK(stage_out(p, x, xd, sizeof(float2) * p.Ntot * p.F));
  return CGS_OK;
  GUARD_END
}

int cgs_normal(cgs_plan plan, const void* x, void* q, double lambda, int normal_op) {
  GUARD_BEGIN
  if (!plan || !x || !q) { set_error("null argument"); return CGS_ERR_ARG; }
  Plan& p = plan->p;
  CHK(require_ready(p, true));
  const float2* xd;
  CHK(stage_in(p, x, (float2*)p.ext_b, sizeof(float2) * p.Ntot * p.F, &xd));
  float2* qd = (p.flags & CGS_DEVICE_IO) ? (float2*)q : p.img[5];
  if (!(p.flags & CGS_DEVICE_IO)) {
    HIPCHK(hipMemcpyAsync(p.img[4], xd, sizeof(float2) * p.Ntot * p.F, hipMemcpyDeviceToDevice, p.stream));
    xd = p.img[4];
  }
  if (normal_op != CGS_NORMAL_NUFFT && normal_op != CGS_NORMAL_TOEPLITZ) { set_error("bad normal operator"); return CGS_ERR_ARG; }
  if (normal_op == CGS_NORMAL_TOEPLITZ) CHK(toeplitz_prepare(p));
  CHK(op_normal(p, normal_op, xd, qd, (float)lambda, nullptr, nullptr));
  CHK(check_launch());
  CHK(stage_out(p, q, qd, sizeof(float2) * p.Ntot * p.F));
  return CGS_OK;
  GUARD_END
}

int cgs_nfft_forward(cgs_plan plan, int frame, const void* img, void* f, int batch) {
  GUARD_BEGIN
  if (!plan || !img || !f) { set_error("null argument"); return CGS_ERR_ARG; }
  Plan& p = plan->p;
  if (frame < 0 || frame >= p.F || batch < 1 || batch > p.C) { set_error("bad frame or batch"); return CGS_ERR_ARG; }
  if (!p.have_k) { set_error("trajectory not set"); return CGS_ERR_STATE; }
  const int64_t Mf = p.foff[frame + 1] - p.foff[frame];
  const float2* id;
  CHK(stage_in(p, img, (float2*)p.ext_b, sizeof(float2) * p.Ntot * batch, &id));
  float2* fd = (p.flags & CGS_DEVICE_IO) ? (float2*)f : (float2*)p.ext_a;
  mask_one_frame(p, frame);
  launch_img_ext2coil(p, id, p.Xc, batch, frame, p.stream);
  if (p.b0q) {   // off-resonance model: samples in the helper plan's order
    CHK(b0_forward(p, p.Xc, p.b0q->ybuf[1], false));
    launch_y_int2ext(*p.b0q, p.b0q->ybuf[1], fd, batch, frame, p.stream);
  } else {
    CHK(fft_forward(p, 2, p.Xc, p.fmask));
    CHK(interp(p, p.ybuf[1], p.fmask));
    launch_y_int2ext(p, p.ybuf[1], fd, batch, frame, p.stream);
  }
  CHK(check_launch());
  CHK(stage_out(p, f, fd, sizeof(float2) * Mf * batch));
  return CGS_OK;
  GUARD_END
}

int cgs_nfft_adjoint(cgs_plan plan, int frame, const void* f, void* img, int batch) {
  GUARD_BEGIN
  if (!plan || !img || !f) { set_error("null argument"); return CGS_ERR_ARG; }
  Plan& p = plan->p;
  if (frame < 0 || frame >= p.F || batch < 1 || batch > p.C) { set_error("bad frame or batch"); return CGS_ERR_ARG; }
  if (!p.have_k) { set_error("trajectory not set"); return CGS_ERR_STATE; }
  const int64_t Mf = p.foff[frame + 1] - p.foff[frame];
  const float2* fd;
  CHK(stage_in(p, f, (float2*)p.ext_a, sizeof(float2) * Mf * batch, &fd));
  float2* id = (p.flags & CGS_DEVICE_IO) ? (float2*)img : (float2*)p.ext_b;
  mask_one_frame(p, frame);
  if (p.b0q) {
    launch_y_ext2int(*p.b0q, fd, p.b0q->ybuf[1], batch, frame, p.stream);
    CHK(b0_adjoint(p, p.b0q->ybuf[1], false, p.Xc));
  } else {
    launch_y_ext2int(p, fd, p.ybuf[1], batch, frame, p.stream);
    CHK(gridding(p, p.ybuf[1], false, p.fmask));
    CHK(fft_adjoint(p, p.Xc, p.fmask));
  }
  launch_img_coil2ext(p, p.Xc, id, batch, frame, p.stream);
  CHK(check_launch());
  CHK(stage_out(p, img, id, sizeof(float2) * p.Ntot * batch));
  return CGS_OK;
  GUARD_END
}

int cgs_nfft_inverse_ref(cgs_plan plan, int frame, const void* f, void* img, int batch, int maxit, double eps) {
  GUARD_BEGIN
  if (!plan || !img || !f) { set_error("null argument"); return CGS_ERR_ARG; }
  Plan& p = plan->p;
  if (frame < 0 || frame >= p.F || batch < 1 || batch > p.C || maxit < 0) { set_error("bad frame, batch or maxit"); return CGS_ERR_ARG; }
  CHK(require_ready(p, false));
  const int64_t Mf = p.foff[frame + 1] - p.foff[frame];
  const float2* fd;
  CHK(stage_in(p, f, (float2*)p.ext_a, sizeof(float2) * Mf * batch, &fd));
  float2* id = (p.flags & CGS_DEVICE_IO) ? (float2*)img : (float2*)p.ext_b;
  mask_one_frame(p, frame);
  if (p.b0q) {
    // the CGNR on the off-resonance model: its sample vectors in the helper's
    // order, so the sample reductions read the helper's weights and frames
    Plan& q = *p.b0q;
    launch_y_ext2int(q, fd, p.ybuf[0], batch, frame, p.stream);
    CnrOps ops;
    ops.adj_w = [&](const float2* r) { return b0_adjoint(p, r, true, p.Xc); };
    ops.fwd_masked = [&](const float2* x, float2* v) { return b0_forward(p, x, v, true); };
    float* w_keep = p.w;
    int* sframe_keep = p.sframe;
    p.w = q.w;
    p.sframe = q.sframe;
    const int rc = cgnr(p, p.ybuf[0], maxit, eps, p.fmask, &ops);
    p.w = w_keep;
    p.sframe = sframe_keep;
    CHK(rc);
  } else {
    launch_y_ext2int(p, fd, p.ybuf[0], batch, frame, p.stream);
    CHK(cgnr(p, p.ybuf[0], maxit, eps, p.fmask));
  }
  launch_img_coil2ext(p, p.Xc2, id, batch, frame, p.stream);
  CHK(check_launch());
  CHK(stage_out(p, img, id, sizeof(float2) * p.Ntot * batch));
  return CGS_OK;
  GUARD_END
}

}  // extern "C"

namespace cgs {

// ---- allocation-free solve ---------------------------------------------------------------
// The CG bookkeeping of cgs_solve_noalloc (per-frame CgState, stop flags,
// reduction partials, residual history) lives in this array: one copy per GPU,
// held until the process ends.  Host code never passes `scratch` itself (in
// host code the name is a stand-in, not a device address): it takes the
// device address once per device with hipGetSymbolAddress and uses that.
static __device__ __align__(256) unsigned char scratch[CGS_NOALLOC_SCRATCH_BYTES];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static unsigned char* g_scratch_dev[kMaxDevices] = {};

static int scratch_address(int dev, unsigned char** out) {
  if (dev < 0 || dev >= kMaxDevices) { set_error("device index out of range"); return CGS_ERR_ARG; }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    int cur = 0;
    HIPCHK(hipGetDevice(&cur));
    if (cur != dev) HIPCHK(hipSetDevice(dev));
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (e != hipSuccess || !ptr) {
      set_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
      return CGS_ERR_HIP;
    }
    g_scratch_dev[dev] = (unsigned char*)ptr;
  }
  *out = g_scratch_dev[dev];
  return CGS_OK;
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// bytes of scratch a solve of this plan with `maxit` iterations needs
static size_t noalloc_bytes(const Plan& p, int maxit) {
  return align256(sizeof(CgState) * p.F) + align256(sizeof(int) * p.F) + align256(sizeof(double) * p.n_partials) +
         sizeof(double) * (size_t)p.F * (size_t)(maxit + 1);
}

// points the plan's CG bookkeeping at the scratch for one call
struct ScratchSwap {
  Plan& p;
  CgState* cgst;
  int* fdone;
  double* partials;
  double* hist;
  int hist_cap;
  ScratchSwap(Plan& p_, unsigned char* base, int hcap)
      : p(p_), cgst(p_.cgst), fdone(p_.fdone), partials(p_.partials), hist(p_.hist), hist_cap(p_.hist_cap) {
    size_t off = 0;
    p.cgst = (CgState*)(base + off);
    off += align256(sizeof(CgState) * p.F);
    p.fdone = (int*)(base + off);
    off += align256(sizeof(int) * p.F);
    p.partials = (double*)(base + off);
    off += align256(sizeof(double) * p.n_partials);
    p.hist = (double*)(base + off);
    p.hist_cap = hcap;
    p.no_prof = true;
  }
  ~ScratchSwap() {
    p.cgst = cgst;
    p.fdone = fdone;
    p.partials = partials;
    p.hist = hist;
    p.hist_cap = hist_cap;
    p.no_prof = false;
  }
};

static int check_solve_opts(const cgs_solve_opts& o) {
  if (o.maxit < 0 || o.ftiter < 0 || !(o.tol >= 0)) { set_error("bad solver options"); return CGS_ERR_ARG; }
  if (o.eh_mode != CGS_EH_TRUE && o.eh_mode != CGS_EH_REF) { set_error("bad E^H mode"); return CGS_ERR_ARG; }
  if (o.normal_op != CGS_NORMAL_NUFFT && o.normal_op != CGS_NORMAL_TOEPLITZ) { set_error("bad normal operator"); return CGS_ERR_ARG; }
  if (o.normal_op == CGS_NORMAL_TOEPLITZ && o.eh_mode != CGS_EH_TRUE) {
    set_error("the Toeplitz normal operator exists for the true adjoint only");
    return CGS_ERR_ARG;
  }
  return CGS_OK;
}

}  // namespace cgs

extern "C" {

int cgs_solve_noalloc(cgs_plan plan, const void* y, const void* x0, void* x, const cgs_solve_opts* opts,
                      cgs_solve_report* report) {
  GUARD_BEGIN
  const auto t_start = std::chrono::steady_clock::now();
  if (!plan || !y || !x || !opts) { set_error("null argument"); return CGS_ERR_ARG; }
  Plan& p = plan->p;
  CHK(require_ready(p, true));
  const cgs_solve_opts o = *opts;
  CHK(check_solve_opts(o));
  if (!(p.flags & CGS_DEVICE_IO)) {
    set_error("cgs_solve_noalloc needs a CGS_DEVICE_IO plan (host buffers would need staging copies)");
    return CGS_ERR_ARG;
  }
  if (o.graph_iters != 0) {
    set_error("cgs_solve_noalloc launches kernels one by one: graph_iters must be 0 (a graph allocates)");
    return CGS_ERR_ARG;
  }
  if (o.eh_mode == CGS_EH_TRUE && o.normal_op == CGS_NORMAL_TOEPLITZ && !p.have_toep) {
    set_error("the Toeplitz spectrum is not prepared (it allocates): call cgs_normal or cgs_solve with "
              "CGS_NORMAL_TOEPLITZ once first");
    return CGS_ERR_STATE;
  }
  const size_t need = noalloc_bytes(p, o.maxit);
  if (need > (size_t)CGS_NOALLOC_SCRATCH_BYTES) {
    set_error("cgs_solve_noalloc: " + std::to_string(need) + " bytes of CG bookkeeping (frames x (maxit + 1) "
              "history, partials) exceed the " + std::to_string((size_t)CGS_NOALLOC_SCRATCH_BYTES) +
              "-byte scratch");
    return CGS_ERR_ARG;
  }
  unsigned char* base = nullptr;
  CHK(scratch_address(p.device, &base));
  const float2* yd = (const float2*)y;
  float2* xd = (float2*)x;
  if (o.maxit == 0) {   // CGLS.hpp:32-33: return A/y
    ScratchSwap sw(p, base, 1);
    launch_y_ext2int(p, yd, p.ybuf[0], p.C, -1, p.stream);
    CHK(op_EH(p, o.eh_mode, p.ybuf[0], xd, o.ftiter, o.fteps, nullptr, 0.f, nullptr, nullptr));
    if (report) { std::memset(report, 0, sizeof(*report)); report->nframes = p.F; }
  } else {
    ScratchSwap sw(p, base, o.maxit + 1);
    CHK(solve(p, yd, (const float2*)x0, xd, o, report, t_start));
  }
  CHK(check_launch());
  HIPCHK(hipStreamSynchronize(p.stream));
  if (report) report->total_ms = ms_since(t_start);
  return (report && report->status == CGS_ERR_NAN) ? CGS_ERR_NAN : CGS_OK;
  GUARD_END
}

int cgs_solve(cgs_plan plan, const void* y, const void* x0, void* x, const cgs_solve_opts* opts,
              cgs_solve_report* report) {
  GUARD_BEGIN
  const auto t_start = std::chrono::steady_clock::now();
  if (!plan || !y || !x || !opts) { set_error("null argument"); return CGS_ERR_ARG; }
  Plan& p = plan->p;
  CHK(require_ready(p, true));
  cgs_solve_opts o = *opts;
  CHK(check_solve_opts(o));
  const int hcap = o.maxit + 1;
  if (hcap > p.hist_cap) {
    // a captured CG graph writes the history through the old pointer and
    // capacity: drop it with the buffer it refers to
    HIPCHK(hipStreamSynchronize(p.stream));
    if (p.graph_exec) (void)hipGraphExecDestroy(p.graph_exec);
    p.graph_exec = nullptr;
    p.graph_key.clear();
    if (p.last_hist == p.hist) { p.last_hist = nullptr; p.last_hist_cap = 0; }
    dfree(p.hist);
    CHK(dalloc(&p.hist, (size_t)p.F * hcap));
    p.hist_cap = hcap;
  }
  const size_t ib = sizeof(float2) * (size_t)p.Ntot * p.F;
  const float2* yd;
  CHK(stage_in(p, y, (float2*)p.ext_a, sizeof(float2) * p.Mtot * p.C, &yd));
  const float2* x0d = nullptr;
  if (x0) {
    if (p.flags & CGS_DEVICE_IO) x0d = (const float2*)x0;
    else {
      HIPCHK(hipMemcpyAsync(p.img[4], x0, ib, hipMemcpyHostToDevice, p.stream));
      x0d = p.img[4];
    }
  }
  float2* xd = (p.flags & CGS_DEVICE_IO) ? (float2*)x : (float2*)p.ext_b;
  if (o.maxit == 0) {   // CGLS.hpp:32-33: return A/y
    Prof pr(p, CGS_K_OTHER, 2.0 * sizeof(float2) * p.Mtot * p.C);
    launch_y_ext2int(p, yd, p.ybuf[0], p.C, -1, p.stream);
    CHK(op_EH(p, o.eh_mode, p.ybuf[0], xd, o.ftiter, o.fteps, nullptr, 0.f, nullptr, nullptr));
    if (report) { std::memset(report, 0, sizeof(*report)); report->nframes = p.F; }
  } else {
    CHK(solve(p, yd, x0d, xd, o, report, t_start));
  }
  CHK(check_launch());
  CHK(stage_out(p, x, xd, ib));
  if (report) report->total_ms = ms_since(t_start);
  return (report && report->status == CGS_ERR_NAN) ? CGS_ERR_NAN : CGS_OK;
  GUARD_END
}

int cgs_profile_mask(cgs_plan plan, unsigned mask) {
  GUARD_BEGIN
  if (!plan) { set_error("null argument"); return CGS_ERR_ARG; }
  plan->p.prof_mask = mask;
  return CGS_OK;
  GUARD_END
}

int cgs_kernel_stats(cgs_plan plan, int kernel, int64_t* launches, double* ms, double* bytes, int reset) {
  GUARD_BEGIN
  if (!plan || kernel < 0 || kernel >= CGS_K_COUNT) { set_error("bad kernel id"); return CGS_ERR_ARG; }
  Plan& p = plan->p;
  CHK(drain_profile(p));
  if (launches) *launches = p.stat[kernel].launches;
  if (ms) *ms = p.stat[kernel].ms;
  if (bytes) *bytes = p.stat[kernel].bytes;
  if (reset) p.stat[kernel] = Plan::Stat{};
  return CGS_OK;
  GUARD_END
}

int cgs_residual_history(cgs_plan plan, int frame, double* out, int cap) {
  GUARD_BEGIN
  if (!plan || !out || cap < 0) { set_error("null argument"); return -CGS_ERR_ARG; }
  Plan& p = plan->p;
  if (frame < 0 || frame >= p.F || !p.last_hist) { set_error("no history for this frame"); return -CGS_ERR_STATE; }
  HIPCHK(hipStreamSynchronize(p.stream));
  // the iteration count of the last solve (its report copy; the device state
  // of an allocation-free solve lives in the shared scratch)
  const int it = p.h_cgst ? p.h_cgst[frame].it : 0;
  const int n = std::min(cap, std::min(it + 1, p.last_hist_cap));
  if (n > 0)
    HIPCHK(hipMemcpy(out, p.last_hist + (size_t)frame * p.last_hist_cap, sizeof(double) * n, hipMemcpyDeviceToHost));
  return n;
  GUARD_END
}

int cgs_plan_info(cgs_plan plan, int* n, int* tile, int64_t* visits, int* grid_items) {
  GUARD_BEGIN
  if (!plan) return CGS_ERR_ARG;
  Plan& p = plan->p;
  if (n) for (int a = 0; a < 3; ++a) n[a] = p.n[a];
  if (tile) *tile = p.T;
  if (visits) *visits = p.n_visits;
  if (grid_items) *grid_items = p.n_grid_items;
  return CGS_OK;
  GUARD_END
}

int cgs_time_toeplitz_conv(cgs_plan plan, int reps, double* us_per_launch) {
  GUARD_BEGIN
  if (!plan || !us_per_launch || reps < 1 || reps > 100000) { set_error("bad argument"); return CGS_ERR_ARG; }
  Plan& p = plan->p;
  if (p.d != 2 || !p.have_toep || !toep_lean_enabled() || !toep_conv_supported((int)p.N[0], (int)p.N[1], p.C)) {
    set_error("needs a 2D plan whose Toeplitz spectrum is prepared, on the lean conv kernel");
    return CGS_ERR_STATE;
  }
  // a scratch copy of the conv's in-place buffer (zeros): the plan's own
  // buffers and CG state are not touched
  const size_t n = (size_t)p.C * 2 * p.N[0] * p.N[1] * p.F;
  float2* buf = nullptr;
  CHK(dalloc(&buf, n));
  launch_zero(buf, sizeof(float2) * n, p.stream);
  ToepArgs a{};
  a.C = p.C; a.N0 = (int)p.N[0]; a.N1 = (int)p.N[1]; a.F = p.F; a.Ntot = p.Ntot;
  a.S = p.S; a.Sfs = p.Sfs; a.fdone = nullptr; a.row = 0;
  a.tw = p.ft[1].tw; a.out = buf; a.That = p.That;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t he = hipEventCreate(&e0);
  if (he == hipSuccess) he = hipEventCreate(&e1);
  float ms = 0.f;
  if (he == hipSuccess) {
    launch_toep_conv(a, p.stream);   // warm-up
    (void)hipEventRecord(e0, p.stream);
    for (int k = 0; k < reps; ++k) launch_toep_conv(a, p.stream);
    (void)hipEventRecord(e1, p.stream);
    he = hipEventSynchronize(e1);
    if (he == hipSuccess) he = hipEventElapsedTime(&ms, e0, e1);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  dfree(buf);
  if (he != hipSuccess) { set_error(std::string("timing failed: ") + hipGetErrorString(he)); return CGS_ERR_HIP; }
  *us_per_launch = 1e3 * (double)ms / reps;
  return CGS_OK;
  GUARD_END
}

int cgs_gridding_kind(cgs_plan plan) {
  if (!plan) return -CGS_ERR_ARG;
  const Plan& p = plan->p;
  if (p.use_grid_bin && p.bin_ids) return CGS_GRID_BINNED;
  if (p.d == 2 && p.use_grid_sweep && p.sw_items) return CGS_GRID_SWEEP_LINES;
  return CGS_GRID_TILES;
}

}  // extern "C"
