/*
 * cgs_oracle.c -- FP64 CPU restatement of codeare's NFFT / NCSENSE / CGLS path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library and the timed CPU
 * port in bench.py.  Never linked into the product.  See cgs_oracle.h for the
 * reference functions each routine follows.
 */
#include "cgs_oracle.h"

#include <math.h>
#include <time.h>
#include <stdlib.h>
#include <string.h>
#include <complex.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

static int g_threads = 0;

/* wall-clock marks of the last ora_cg_true / ora_cgls_ref call (the CPU-baseline
 * leg times them): t[0] after E^H y, t[i] after iteration i */
#define ORA_TMARKS 4096
static double g_tmark[ORA_TMARKS + 1];
static int g_ntmark = 0;
static double g_tstart = 0.0;
static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}
static void tmark_reset(void) { g_ntmark = 0; g_tstart = now_s(); }
static void tmark(void) {
  if (g_ntmark <= ORA_TMARKS) g_tmark[g_ntmark++] = now_s() - g_tstart;
}
int ora_time_marks(double* out, int cap) {
  const int n = g_ntmark < cap ? g_ntmark : cap;
  for (int i = 0; i < n; ++i) out[i] = g_tmark[i];
  return n;
}
void ora_set_threads(int n) { g_threads = n; }
int ora_get_threads(void) {
#ifdef _OPENMP
  return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------------- */
/* Bessel I0 by its power series sum_k (x/2)^(2k)/(k!)^2 (all terms positive). */
double ora_bessel_i0(double x) {
  double q = 0.25 * x * x, term = 1.0, sum = 1.0;
  for (int k = 1; k < 500; ++k) {
    term *= q / ((double)k * (double)k);
    sum += term;
    if (term < 1e-17 * sum) break;
  }
  return sum;
}

/* NFFTTraits.hpp:665-673 transcribes NFFT3's KB window; at |t| == m the limit
 * of sinh(b s)/(pi s) as s->0 is b/pi (the transcription writes 1.0 there). */
double ora_kb_phi(double t, int m, double b) {
  double s2 = (double)m * m - t * t;
  if (s2 > 0) { double s = sqrt(s2); return sinh(b * s) / (M_PI * s); }
  if (s2 < 0) { double s = sqrt(-s2); return sin(b * s) / (M_PI * s); }
  return b / M_PI;
}

/* NFFTTraits.hpp:660-663 (phi_hut). */
double ora_kb_phihut(double k, int n, int m, double b) {
  double a = 2.0 * M_PI * k / n;
  double s2 = b * b - a * a;
  return ora_bessel_i0(m * sqrt(s2 > 0 ? s2 : 0.0));
}

/* ------------------------------------------------------------------------- */
/* Mixed-radix Stockham FFT (textbook autosort form).  Radices 4,2,3,5 and a
 * direct DFT for any other prime factor. */
typedef struct { int n, nf, f[32]; cplx* w; } fftplan;

static fftplan* fft_plan(int n) {
  fftplan* P = (fftplan*)calloc(1, sizeof(fftplan));
  P->n = n;
  int r = n;
  while (r % 4 == 0) { P->f[P->nf++] = 4; r /= 4; }
  while (r % 2 == 0) { P->f[P->nf++] = 2; r /= 2; }
  for (int q = 3; r > 1; q += 2)
    while (r % q == 0) { P->f[P->nf++] = q; r /= q; }
  P->w = (cplx*)malloc(sizeof(cplx) * (size_t)n);
  for (int p = 0; p < n; ++p) {
    double a = -2.0 * M_PI * (double)p / (double)n;
    P->w[p] = cos(a) + I * sin(a);
  }
  return P;
}
static void fft_free(fftplan* P) { if (P) { free(P->w); free(P); } }

static inline cplx tw(const fftplan* P, long p, int sign) {
  cplx v = P->w[p % P->n];
  return sign < 0 ? v : conj(v);
}

/* x: input (n), y: scratch (n).  Result is written back to x. */
static void fft_exec(const fftplan* P, cplx* x, cplx* y, int sign) {
  int n = P->n;
  cplx *src = x, *dst = y;
  int Ns = 1;
  cplx v[64], o[64];
  for (int s = 0; s < P->nf; ++s) {
    int R = P->f[s], nR = n / R;
    long stride = (long)n / ((long)Ns * R);
    for (int j = 0; j < nR; ++j) {
      int k = j % Ns;
      for (int r = 0; r < R; ++r) v[r] = src[j + r * nR] * tw(P, (long)r * k * stride, sign);
      for (int q = 0; q < R; ++q) {
        cplx acc = 0;
        for (int r = 0; r < R; ++r) acc += v[r] * tw(P, (long)((r * q) % R) * (n / R), sign);
        o[q] = acc;
      }
      int idxD = (j / Ns) * Ns * R + k;
      for (int q = 0; q < R; ++q) dst[idxD + q * Ns] = o[q];
    }
    cplx* t = src; src = dst; dst = t;
    Ns *= R;
  }
  if (src != x) memcpy(x, src, sizeof(cplx) * (size_t)n);
}

int ora_fft1d(double* x, int n, int sign) {
  fftplan* P = fft_plan(n);
  cplx* y = (cplx*)malloc(sizeof(cplx) * (size_t)n);
  fft_exec(P, (cplx*)x, y, sign);
  free(y);
  fft_free(P);
  return 0;
}

/* n-D FFT of a column-major array (axis 0 fastest). */
static void fft_nd(fftplan* const* P, int d, const int* n, cplx* a, int sign) {
  int nmax = 0;
  for (int t = 0; t < d; ++t) if (n[t] > nmax) nmax = n[t];
  cplx* line = (cplx*)malloc(sizeof(cplx) * (size_t)nmax);
  cplx* scr = (cplx*)malloc(sizeof(cplx) * (size_t)nmax);
  long tot = 1;
  for (int t = 0; t < d; ++t) tot *= n[t];
  long stride = 1;
  for (int ax = 0; ax < d; ++ax) {
    long len = n[ax], nlines = tot / len;
    for (long L = 0; L < nlines; ++L) {
      long lo = L % stride, hi = L / stride;
      long base = lo + hi * stride * len;
      for (long p = 0; p < len; ++p) line[p] = a[base + p * stride];
      fft_exec(P[ax], line, scr, sign);
      for (long p = 0; p < len; ++p) a[base + p * stride] = line[p];
    }
    stride *= len;
  }
  free(line);
  free(scr);
}

/* ------------------------------------------------------------------------- */
struct ora_nfft {
  int d, N[3], n[3], m, taps;
  int64_t M, Ntot, ntot;
  double b[3];
  double* dinv[3];  /* 1/phi_hut per image index along each axis */
  int* base;        /* (floor(n x) - m) mod n, per sample per axis */
  double* psi;      /* window values, per sample per axis per tap */
  fftplan* fp[3];
};

ora_nfft* ora_nfft_create(int d, const int* N, int64_t M, double sigma, int m) {
  if (d < 1 || d > 3 || M < 0 || m < 0 || sigma < 1.0) return NULL;
  ora_nfft* p = (ora_nfft*)calloc(1, sizeof(ora_nfft));
  p->d = d; p->M = M; p->m = m; p->taps = 2 * m + 2;
  p->Ntot = 1; p->ntot = 1;
  for (int a = 0; a < d; ++a) {
    p->N[a] = N[a];
    p->n[a] = (int)ceil(sigma * (double)N[a]);   /* NFFT.hpp:108-109 */
    p->Ntot *= N[a]; p->ntot *= p->n[a];
    double s = (double)p->n[a] / (double)N[a];
    p->b[a] = M_PI * (2.0 - 1.0 / s);
    p->dinv[a] = (double*)malloc(sizeof(double) * (size_t)N[a]);
    for (int i = 0; i < N[a]; ++i)
      p->dinv[a][i] = 1.0 / ora_kb_phihut((double)(i - N[a] / 2), p->n[a], m, p->b[a]);
    p->fp[a] = fft_plan(p->n[a]);
  }
  for (int a = d; a < 3; ++a) { p->N[a] = 1; p->n[a] = 1; }
  p->base = (int*)calloc((size_t)M * d, sizeof(int));
  p->psi = (double*)calloc((size_t)M * d * p->taps, sizeof(double));
  return p;
}

void ora_nfft_destroy(ora_nfft* p) {
  if (!p) return;
  for (int a = 0; a < p->d; ++a) { free(p->dinv[a]); fft_free(p->fp[a]); }
  free(p->base); free(p->psi); free(p);
}

int ora_nfft_n(const ora_nfft* p, int axis) { return p->n[axis]; }

/* NFFT::KSpace + NFFTTraits::Psi (NFFT.hpp:258-296): k(t,j) feeds image axis d-1-t. */
void ora_nfft_set_trajectory(ora_nfft* p, const float* k) {
  int d = p->d, T = p->taps, m = p->m;
  for (int64_t j = 0; j < p->M; ++j) {
    for (int a = 0; a < d; ++a) {
      double x = (double)k[(d - 1 - a) + (int64_t)d * j];
      double u = (double)p->n[a] * x;
      double fl = floor(u), fr = u - fl;
      long bs = ((long)fl - m) % p->n[a];
      if (bs < 0) bs += p->n[a];
      p->base[j * d + a] = (int)bs;
      for (int s = 0; s < T; ++s)
        p->psi[(j * d + a) * T + s] = ora_kb_phi(fr + m - s, m, p->b[a]);
    }
  }
}

static inline int64_t grid_of_image(const ora_nfft* p, int64_t i, double* dfac) {
  int64_t g = 0, mul = 1, rem = i;
  double f = 1.0;
  for (int a = 0; a < p->d; ++a) {
    int ia = (int)(rem % p->N[a]);
    rem /= p->N[a];
    long ka = ia - p->N[a] / 2;
    long la = ((ka % p->n[a]) + p->n[a]) % p->n[a];
    g += la * mul;
    mul *= p->n[a];
    f *= p->dinv[a][ia];
  }
  *dfac = f;
  return g;
}

/* apply the window taps of sample j: forward (gather) or adjoint (scatter) */
static inline cplx interp_one(const ora_nfft* p, const cplx* g, int64_t j) {
  int d = p->d, T = p->taps;
  const double* ps = p->psi + (size_t)j * d * T;
  const int* bs = p->base + (size_t)j * d;
  cplx acc = 0;
  if (d == 1) {
    for (int s0 = 0; s0 < T; ++s0) acc += ps[s0] * g[(bs[0] + s0) % p->n[0]];
  } else if (d == 2) {
    for (int s1 = 0; s1 < T; ++s1) {
      int64_t l1 = (bs[1] + s1) % p->n[1];
      cplx row = 0;
      for (int s0 = 0; s0 < T; ++s0) row += ps[s0] * g[(bs[0] + s0) % p->n[0] + p->n[0] * l1];
      acc += ps[T + s1] * row;
    }
  } else {
    for (int s2 = 0; s2 < T; ++s2) {
      int64_t l2 = (bs[2] + s2) % p->n[2];
      for (int s1 = 0; s1 < T; ++s1) {
        int64_t l1 = (bs[1] + s1) % p->n[1];
        cplx row = 0;
        for (int s0 = 0; s0 < T; ++s0)
          row += ps[s0] * g[(bs[0] + s0) % p->n[0] + p->n[0] * (l1 + p->n[1] * l2)];
        acc += ps[2 * T + s2] * ps[T + s1] * row;
      }
    }
  }
  return acc;
}

static inline void spread_one(const ora_nfft* p, cplx* g, int64_t j, cplx v) {
  int d = p->d, T = p->taps;
  const double* ps = p->psi + (size_t)j * d * T;
  const int* bs = p->base + (size_t)j * d;
  if (d == 1) {
    for (int s0 = 0; s0 < T; ++s0) g[(bs[0] + s0) % p->n[0]] += ps[s0] * v;
  } else if (d == 2) {
    for (int s1 = 0; s1 < T; ++s1) {
      int64_t l1 = (bs[1] + s1) % p->n[1];
      cplx vr = ps[T + s1] * v;
      for (int s0 = 0; s0 < T; ++s0) g[(bs[0] + s0) % p->n[0] + p->n[0] * l1] += ps[s0] * vr;
    }
  } else {
    for (int s2 = 0; s2 < T; ++s2) {
      int64_t l2 = (bs[2] + s2) % p->n[2];
      for (int s1 = 0; s1 < T; ++s1) {
        int64_t l1 = (bs[1] + s1) % p->n[1];
        cplx vr = ps[2 * T + s2] * ps[T + s1] * v;
        for (int s0 = 0; s0 < T; ++s0)
          g[(bs[0] + s0) % p->n[0] + p->n[0] * (l1 + p->n[1] * l2)] += ps[s0] * vr;
      }
    }
  }
}

/* nfft_trafo: D (deconvolve + zero-pad), F (sign -1), B (window gather). */
void ora_nfft_trafo(const ora_nfft* p, const double* img_, double* f_) {
  const cplx* img = (const cplx*)img_;
  cplx* f = (cplx*)f_;
  cplx* g = (cplx*)calloc((size_t)p->ntot, sizeof(cplx));
  for (int64_t i = 0; i < p->Ntot; ++i) {
    double df;
    int64_t gi = grid_of_image(p, i, &df);
    g[gi] = img[i] * df;
  }
  fft_nd((fftplan* const*)p->fp, p->d, p->n, g, -1);
  for (int64_t j = 0; j < p->M; ++j) f[j] = interp_one(p, g, j);
  free(g);
}

/* nfft_adjoint: B^T (window scatter), F (sign +1), D (crop + deconvolve). */
void ora_nfft_adjoint(const ora_nfft* p, const double* f_, double* img_) {
  const cplx* f = (const cplx*)f_;
  cplx* img = (cplx*)img_;
  cplx* g = (cplx*)calloc((size_t)p->ntot, sizeof(cplx));
  for (int64_t j = 0; j < p->M; ++j) spread_one(p, g, j, f[j]);
  fft_nd((fftplan* const*)p->fp, p->d, p->n, g, +1);
  for (int64_t i = 0; i < p->Ntot; ++i) {
    double df;
    int64_t gi = grid_of_image(p, i, &df);
    img[i] = g[gi] * df;
  }
  free(g);
}

/* ------------------------------------------------------------------------- */
/* Direct NDFT, separable exponentials per axis. */
static void ndft_core(int d, const int* N, int64_t M, const float* k,
                      const cplx* in, cplx* out, int adjoint) {
  int64_t Ntot = 1;
  for (int a = 0; a < d; ++a) Ntot *= N[a];
  double sg = adjoint ? +1.0 : -1.0;
  if (adjoint) memset(out, 0, sizeof(cplx) * (size_t)Ntot);
  int Nmax = 0;
  for (int a = 0; a < d; ++a) if (N[a] > Nmax) Nmax = N[a];
  cplx* e = (cplx*)malloc(sizeof(cplx) * (size_t)3 * Nmax);
  for (int64_t j = 0; j < M; ++j) {
    for (int a = 0; a < d; ++a) {
      double x = (double)k[(d - 1 - a) + (int64_t)d * j];
      for (int i = 0; i < N[a]; ++i) {
        double ph = sg * 2.0 * M_PI * (double)(i - N[a] / 2) * x;
        e[a * Nmax + i] = cos(ph) + I * sin(ph);
      }
    }
    if (!adjoint) {
      cplx acc = 0;
      for (int64_t i = 0; i < Ntot; ++i) {
        int64_t r = i; cplx ph = 1;
        for (int a = 0; a < d; ++a) { ph *= e[a * Nmax + r % N[a]]; r /= N[a]; }
        acc += in[i] * ph;
      }
      out[j] = acc;
    } else {
      for (int64_t i = 0; i < Ntot; ++i) {
        int64_t r = i; cplx ph = 1;
        for (int a = 0; a < d; ++a) { ph *= e[a * Nmax + r % N[a]]; r /= N[a]; }
        out[i] += in[j] * ph;
      }
    }
  }
  free(e);
}

void ora_ndft_trafo(int d, const int* N, int64_t M, const float* k, const double* img, double* f) {
  ndft_core(d, N, M, k, (const cplx*)img, (cplx*)f, 0);
}
void ora_ndft_adjoint(int d, const int* N, int64_t M, const float* k, const double* f, double* img) {
  ndft_core(d, N, M, k, (const cplx*)f, (cplx*)img, 1);
}

/* ------------------------------------------------------------------------- */
/* NFFTTraits::Weights (NFFTTraits.hpp:225-260): disk/ball of radius N0/2 about
 * N0/2, with N0 = plan.N[0] used for every axis. */
void ora_damping_mask(int d, const int* N, double* what) {
  int n0 = N[0], nh = (int)(0.5 * n0);
  int64_t tot = 1;
  for (int a = 0; a < d; ++a) tot *= n0;
  for (int64_t i = 0; i < tot; ++i) {
    int64_t r = i, s = 0;
    for (int a = 0; a < d; ++a) { int64_t q = r % n0 - nh; s += q * q; r /= n0; }
    what[i] = (sqrt((double)s) > nh) ? 0.0 : 1.0;
  }
}

static double dot_w(const cplx* x, const double* w, int64_t n) {
  double s = 0;
  for (int64_t i = 0; i < n; ++i) s += w[i] * (creal(x[i]) * creal(x[i]) + cimag(x[i]) * cimag(x[i]));
  return s;
}

/* NFFTTraits::ITrafo (NFFTTraits.hpp:122-138) driving NFFT3's CGNR with
 * PRECOMPUTE_WEIGHT | PRECOMPUTE_DAMP, restated from the published solver. */
int ora_itrafo(const ora_nfft* p, const double* y_, const double* w,
               const double* what, int maxit, double eps, double* fhat_, double* hist) {
  int64_t M = p->M, Nt = p->Ntot;
  const cplx* y = (const cplx*)y_;
  cplx* fh = (cplx*)fhat_;
  cplx* r = (cplx*)malloc(sizeof(cplx) * (size_t)M);
  cplx* v = (cplx*)malloc(sizeof(cplx) * (size_t)M);
  cplx* z = (cplx*)malloc(sizeof(cplx) * (size_t)Nt);
  cplx* pp = (cplx*)malloc(sizeof(cplx) * (size_t)Nt);
  cplx* t = (cplx*)malloc(sizeof(cplx) * (size_t)(Nt > M ? Nt : M));
  memset(fh, 0, sizeof(cplx) * (size_t)Nt);                  /* :126 */
  /* solver_before_loop: r = y - A*0 = y */
  for (int64_t j = 0; j < M; ++j) r[j] = y[j];
  double dot_r = dot_w(r, w, M);
  for (int64_t j = 0; j < M; ++j) t[j] = w[j] * r[j];
  ora_nfft_adjoint(p, (const double*)t, (double*)z);
  double dot_z = dot_w(z, what, Nt);
  for (int64_t i = 0; i < Nt; ++i) pp[i] = z[i];
  int it = 0;
  for (int l = 0; l < maxit; ++l) {                          /* :130-134 */
    if (hist) hist[l] = dot_r;
    if (dot_r < eps) break;
    for (int64_t i = 0; i < Nt; ++i) t[i] = what[i] * pp[i];
    ora_nfft_trafo(p, (const double*)t, (double*)v);
    double dot_v = dot_w(v, w, M);
    double alpha = dot_z / dot_v;
    for (int64_t i = 0; i < Nt; ++i) fh[i] += alpha * what[i] * pp[i];
    for (int64_t j = 0; j < M; ++j) r[j] -= alpha * v[j];
    dot_r = dot_w(r, w, M);
    for (int64_t j = 0; j < M; ++j) t[j] = w[j] * r[j];
    ora_nfft_adjoint(p, (const double*)t, (double*)z);
    double dz_old = dot_z;
    dot_z = dot_w(z, what, Nt);
    double beta = dot_z / dz_old;
    for (int64_t i = 0; i < Nt; ++i) pp[i] = beta * pp[i] + z[i];
    ++it;
  }
  if (hist) hist[it] = dot_r;
  free(r); free(v); free(z); free(pp); free(t);
  return it;
}

/* ------------------------------------------------------------------------- */
/* IntensityMap (MRI.hpp:17-28), FP32 accumulation like the reference. */
void ora_intensity_map(int64_t nr, int C, const float* S, float* Imap) {
  for (int64_t i = 0; i < nr; ++i) {
    float acc = 0.0f;
    for (int c = 0; c < C; ++c) {
      float re = S[2 * (i + nr * c)], im = S[2 * (i + nr * c) + 1];
      acc += re * re + im * im;
    }
    Imap[i] = (float)(1.0 / (sqrt((double)acc) + 1.e-9));
  }
}

static int nthreads_for(int C) {
  int t = ora_get_threads();
  return t < C ? t : C;
}

/* NCSENSE::Trafo (NCSENSE.hpp:334-347): y(:,c) = NFFT(S_c .* x).  The product
 * S_c .* x is formed in complex64 like the reference's View*Matrix. */
void ora_ncsense_E(const ora_nfft* p, int C, const float* S, const float* x, float* y) {
  int64_t Nt = p->Ntot, M = p->M;
#pragma omp parallel for schedule(dynamic) num_threads(nthreads_for(C))
  for (int c = 0; c < C; ++c) {
    cplx* in = (cplx*)malloc(sizeof(cplx) * (size_t)Nt);
    cplx* out = (cplx*)malloc(sizeof(cplx) * (size_t)M);
    for (int64_t i = 0; i < Nt; ++i) {
      float sr = S[2 * (i + Nt * c)], si = S[2 * (i + Nt * c) + 1];
      float xr = x[2 * i], xi = x[2 * i + 1];
      float pr = sr * xr - si * xi, pi = sr * xi + si * xr;
      in[i] = (double)pr + I * (double)pi;
    }
    ora_nfft_trafo(p, (const double*)in, (double*)out);
    for (int64_t j = 0; j < M; ++j) {
      y[2 * (j + M * c)] = (float)creal(out[j]);
      y[2 * (j + M * c) + 1] = (float)cimag(out[j]);
    }
    free(in); free(out);
  }
}

/* Density-weighted true adjoint: x = sum_c conj(S_c) .* NFFT^H(w .* y_c). */
void ora_ncsense_EH_true(const ora_nfft* p, int C, const float* S, const float* w,
                         const float* y, float* x) {
  int64_t Nt = p->Ntot, M = p->M;
  cplx* acc = (cplx*)calloc((size_t)Nt * C, sizeof(cplx));
#pragma omp parallel for schedule(dynamic) num_threads(nthreads_for(C))
  for (int c = 0; c < C; ++c) {
    cplx* in = (cplx*)malloc(sizeof(cplx) * (size_t)M);
    cplx* out = acc + (size_t)Nt * c;
    for (int64_t j = 0; j < M; ++j)
      in[j] = (double)w[j] * ((double)y[2 * (j + M * c)] + I * (double)y[2 * (j + M * c) + 1]);
    ora_nfft_adjoint(p, (const double*)in, (double*)out);
    for (int64_t i = 0; i < Nt; ++i) {
      cplx s = (double)S[2 * (i + Nt * c)] + I * (double)S[2 * (i + Nt * c) + 1];
      out[i] *= conj(s);
    }
    free(in);
  }
  for (int64_t i = 0; i < Nt; ++i) {
    cplx s = 0;
    for (int c = 0; c < C; ++c) s += acc[i + Nt * c];
    x[2 * i] = (float)creal(s);
    x[2 * i + 1] = (float)cimag(s);
  }
  free(acc);
}

/* NCSENSE::operator/ (NCSENSE.hpp:289-298): x = I .* sum_c conj(S_c) .* ITrafo_c(y_c). */
void ora_ncsense_EH_ref(const ora_nfft* p, int C, const float* S, const float* w,
                        int ftiter, double fteps, const float* y, float* x) {
  int64_t Nt = p->Ntot, M = p->M;
  double* what = (double*)malloc(sizeof(double) * (size_t)Nt);
  ora_damping_mask(p->d, p->N, what);
  double* wd = (double*)malloc(sizeof(double) * (size_t)M);
  for (int64_t j = 0; j < M; ++j) wd[j] = w[j];
  float* Iw = (float*)malloc(sizeof(float) * (size_t)Nt);
  ora_intensity_map(Nt, C, S, Iw);
  cplx* acc = (cplx*)calloc((size_t)Nt * C, sizeof(cplx));
#pragma omp parallel for schedule(dynamic) num_threads(nthreads_for(C))
  for (int c = 0; c < C; ++c) {
    cplx* in = (cplx*)malloc(sizeof(cplx) * (size_t)M);
    cplx* out = acc + (size_t)Nt * c;
    for (int64_t j = 0; j < M; ++j)
      in[j] = (double)y[2 * (j + M * c)] + I * (double)y[2 * (j + M * c) + 1];
    ora_itrafo(p, (const double*)in, wd, what, ftiter, fteps, (double*)out, NULL);
    /* NFFT::Adjoint copies f_hat_iter to complex64 (NFFT.hpp:399-402) */
    for (int64_t i = 0; i < Nt; ++i) {
      cplx s = (double)S[2 * (i + Nt * c)] + I * (double)S[2 * (i + Nt * c) + 1];
      cplx o = (double)(float)creal(out[i]) + I * (double)(float)cimag(out[i]);
      out[i] = o * conj(s);
    }
    free(in);
  }
  for (int64_t i = 0; i < Nt; ++i) {
    cplx s = 0;
    for (int c = 0; c < C; ++c) s += acc[i + Nt * c];
    s *= (double)Iw[i];
    x[2 * i] = (float)creal(s);
    x[2 * i + 1] = (float)cimag(s);
  }
  free(acc); free(what); free(wd); free(Iw);
}

/* ------------------------------------------------------------------------- */
/* dotc (Lapack.hpp:619-624): serial complex64 accumulate of conj(a)*b. */
static float complex dotc_f(const float complex* a, const float complex* b, int64_t n) {
  float complex s = 0;
  for (int64_t i = 0; i < n; ++i) s += conjf(a[i]) * b[i];
  return s;
}

static void apply_EH(const ora_nfft* p, int C, const float* S, const float* w, int mode,
                     int ftiter, double fteps, const float* y, float* x) {
  if (mode == 0) ora_ncsense_EH_true(p, C, S, w, y, x);
  else ora_ncsense_EH_ref(p, C, S, w, ftiter, fteps, y, x);
}

/* CGLS::Solve (CGLS.hpp:30-59) with complex64 vectors and scalars. */
int ora_cgls_ref(const ora_nfft* p, int C, const float* S, const float* w,
                 int mode, int ftiter, double fteps, int maxit, double cgeps,
                 double lambda, const float* y, float* x_, double* hist) {
  int64_t Nt = p->Ntot, M = p->M;
  float complex* x = (float complex*)x_;
  float complex* pp = (float complex*)malloc(sizeof(float complex) * (size_t)Nt);
  float complex* r = (float complex*)malloc(sizeof(float complex) * (size_t)Nt);
  float complex* q = (float complex*)malloc(sizeof(float complex) * (size_t)Nt);
  float* yt = (float*)malloc(sizeof(float) * 2 * (size_t)M * C);
  tmark_reset();
  apply_EH(p, C, S, w, mode, ftiter, fteps, y, (float*)pp);        /* :31 */
  tmark();
  if (maxit == 0) {                                                /* :32-33 */
    memcpy(x, pp, sizeof(float complex) * (size_t)Nt);
    free(pp); free(r); free(q); free(yt);
    return 0;
  }
  memcpy(r, pp, sizeof(float complex) * (size_t)Nt);
  float rn = crealf(dotc_f(r, pp, Nt)), xn = rn, rno;               /* :35 */
  memset(x, 0, sizeof(float complex) * (size_t)Nt);
  int it = 0;
  for (int i = 0; i < maxit; ++i) {
    float res = rn / xn;
    if (hist) hist[i] = res;
    if (isnan(res) || res <= (float)cgeps) break;                  /* :41 */
    ora_ncsense_E(p, C, S, (const float*)pp, yt);                  /* :47 */
    apply_EH(p, C, S, w, mode, ftiter, fteps, yt, (float*)q);
    if (lambda != 0.0)
      for (int64_t k = 0; k < Nt; ++k) q[k] += (float)lambda * pp[k];
    float ts = rn / crealf(dotc_f(pp, q, Nt));                     /* :50 */
    for (int64_t k = 0; k < Nt; ++k) x[k] += ts * pp[k];
    for (int64_t k = 0; k < Nt; ++k) r[k] -= ts * q[k];
    rno = rn;
    rn = crealf(dotc_f(r, r, Nt));
    for (int64_t k = 0; k < Nt; ++k) pp[k] = pp[k] * (rn / rno) + r[k];   /* :55-56 */
    ++it;
    tmark();
  }
  if (hist && it == maxit) hist[it] = rn / xn;
  free(pp); free(r); free(q); free(yt);
  return it;
}

/* ------------------------------------------------------------------------- */
/* Frames (NCSENSE nmany, NCSENSE.hpp:141-151, :254-288, :311-333) and Cartesian
 * slices (3rd_dim_cart, NCSENSE.hpp:81-99, :338-339): frame f has its own plan
 * (trajectory), maps S + 2*f*Sfs floats, weights w + moff[f] and data block
 * y + 2*C*moff[f] of shape (M_f, C); the images are concatenated (N^d, F). */
static void frames_E(const ora_nfft* const* pl, int F, int C, const float* S, int64_t Sfs,
                     const int64_t* moff, const float* x, float* y) {
  const int64_t Nt = pl[0]->Ntot;
  for (int f = 0; f < F; ++f)
    ora_ncsense_E(pl[f], C, S + 2 * f * Sfs, x + 2 * f * Nt, y + 2 * (int64_t)C * moff[f]);
}

static void frames_EH(const ora_nfft* const* pl, int F, int C, const float* S, int64_t Sfs,
                      const float* w, const int64_t* moff, int mode, int ftiter, double fteps,
                      const float* y, float* x) {
  const int64_t Nt = pl[0]->Ntot;
  for (int f = 0; f < F; ++f)
    apply_EH(pl[f], C, S + 2 * f * Sfs, w + moff[f], mode, ftiter, fteps, y + 2 * (int64_t)C * moff[f],
             x + 2 * f * Nt);
}

/* CGLS::Solve (CGLS.hpp:30-59) on the whole (N^d, F) image matrix, as codeare
 * runs it for nmany frames and 3rd_dim_cart slices: one alpha and beta from
 * serial complex64 dotc over the concatenated vectors (Lapack.hpp:619-624). */
int ora_cgls_ref_frames(const ora_nfft* const* pl, int F, int C, const float* S, int64_t Sfs,
                        const float* w, const int64_t* moff, int mode, int ftiter, double fteps,
                        int maxit, double cgeps, double lambda, const float* y, float* x_, double* hist) {
  const int64_t Nt = pl[0]->Ntot * F, Mt = moff[F];
  float complex* x = (float complex*)x_;
  float complex* pp = (float complex*)malloc(sizeof(float complex) * (size_t)Nt);
  float complex* r = (float complex*)malloc(sizeof(float complex) * (size_t)Nt);
  float complex* q = (float complex*)malloc(sizeof(float complex) * (size_t)Nt);
  float* yt = (float*)malloc(sizeof(float) * 2 * (size_t)Mt * C);
  frames_EH(pl, F, C, S, Sfs, w, moff, mode, ftiter, fteps, y, (float*)pp);   /* :31 */
  if (maxit == 0) {                                                          /* :32-33 */
    memcpy(x, pp, sizeof(float complex) * (size_t)Nt);
    free(pp); free(r); free(q); free(yt);
    return 0;
  }
  memcpy(r, pp, sizeof(float complex) * (size_t)Nt);
  float rn = crealf(dotc_f(r, pp, Nt)), xn = rn, rno;                         /* :35 */
  memset(x, 0, sizeof(float complex) * (size_t)Nt);
  int it = 0;
  for (int i = 0; i < maxit; ++i) {
    float res = rn / xn;
    if (hist) hist[i] = res;
    if (isnan(res) || res <= (float)cgeps) break;                            /* :41 */
    frames_E(pl, F, C, S, Sfs, moff, (const float*)pp, yt);                  /* :47 */
    frames_EH(pl, F, C, S, Sfs, w, moff, mode, ftiter, fteps, yt, (float*)q);
    if (lambda != 0.0)
      for (int64_t k = 0; k < Nt; ++k) q[k] += (float)lambda * pp[k];
    float ts = rn / crealf(dotc_f(pp, q, Nt));                               /* :50 */
    for (int64_t k = 0; k < Nt; ++k) x[k] += ts * pp[k];
    for (int64_t k = 0; k < Nt; ++k) r[k] -= ts * q[k];
    rno = rn;
    rn = crealf(dotc_f(r, r, Nt));
    for (int64_t k = 0; k < Nt; ++k) pp[k] = pp[k] * (rn / rno) + r[k];      /* :55-56 */
    ++it;
  }
  if (hist && it == maxit) hist[it] = rn / xn;
  free(pp); free(r); free(q); free(yt);
  return it;
}

/* ora_cg_true on the concatenated frames with one set of FP64 scalars (the GPU
 * library's joint solve). */
int ora_cg_true_frames(const ora_nfft* const* pl, int F, int C, const float* S, int64_t Sfs,
                       const float* w, const int64_t* moff, int maxit, double tol, double lambda,
                       const float* y, float* x_, double* hist) {
  const int64_t Nt = pl[0]->Ntot * F, Mt = moff[F];
  float complex* x = (float complex*)x_;
  float complex* pp = (float complex*)malloc(sizeof(float complex) * (size_t)Nt);
  float complex* r = (float complex*)malloc(sizeof(float complex) * (size_t)Nt);
  float complex* q = (float complex*)malloc(sizeof(float complex) * (size_t)Nt);
  float* yt = (float*)malloc(sizeof(float) * 2 * (size_t)Mt * C);
  frames_EH(pl, F, C, S, Sfs, w, moff, 0, 0, 0.0, y, (float*)r);
  double bn = 0;
  for (int64_t k = 0; k < Nt; ++k) bn += (double)crealf(r[k]) * crealf(r[k]) + (double)cimagf(r[k]) * cimagf(r[k]);
  memset(x, 0, sizeof(float complex) * (size_t)Nt);
  memcpy(pp, r, sizeof(float complex) * (size_t)Nt);
  double rn = bn;
  int it = 0;
  for (;;) {
    double rel = bn > 0 ? sqrt(rn / bn) : 0.0;
    if (hist) hist[it] = rel;
    if (isnan(rel) || rel <= tol || it >= maxit) break;
    frames_E(pl, F, C, S, Sfs, moff, (const float*)pp, yt);
    frames_EH(pl, F, C, S, Sfs, w, moff, 0, 0, 0.0, yt, (float*)q);
    double pq = 0;
    for (int64_t k = 0; k < Nt; ++k) {
      q[k] += (float)lambda * pp[k];
      pq += (double)crealf(pp[k]) * crealf(q[k]) + (double)cimagf(pp[k]) * cimagf(q[k]);
    }
    double alpha = rn / pq;
    double rn_new = 0;
    for (int64_t k = 0; k < Nt; ++k) {
      x[k] += (float)alpha * pp[k];
      r[k] -= (float)alpha * q[k];
      rn_new += (double)crealf(r[k]) * crealf(r[k]) + (double)cimagf(r[k]) * cimagf(r[k]);
    }
    double beta = rn_new / rn;
    rn = rn_new;
    for (int64_t k = 0; k < Nt; ++k) pp[k] = r[k] + (float)beta * pp[k];
    ++it;
  }
  free(pp); free(r); free(q); free(yt);
  return it;
}

/* CG on the true normal equations with FP64 scalars (the GPU solver's algorithm). */
int ora_cg_true(const ora_nfft* p, int C, const float* S, const float* w,
                int maxit, double tol, double lambda, const float* y,
                const float* x0, float* x_, double* hist) {
  int64_t Nt = p->Ntot, M = p->M;
  float complex* x = (float complex*)x_;
  float complex* b = (float complex*)malloc(sizeof(float complex) * (size_t)Nt);
  float complex* pp = (float complex*)malloc(sizeof(float complex) * (size_t)Nt);
  float complex* r = (float complex*)malloc(sizeof(float complex) * (size_t)Nt);
  float complex* q = (float complex*)malloc(sizeof(float complex) * (size_t)Nt);
  float* yt = (float*)malloc(sizeof(float) * 2 * (size_t)M * C);
  tmark_reset();
  ora_ncsense_EH_true(p, C, S, w, y, (float*)b);
  tmark();
  double bn = 0;
  for (int64_t k = 0; k < Nt; ++k) bn += (double)crealf(b[k]) * crealf(b[k]) + (double)cimagf(b[k]) * cimagf(b[k]);
  if (x0) {
    memcpy(x, x0, sizeof(float complex) * (size_t)Nt);
    ora_ncsense_E(p, C, S, (const float*)x, yt);
    ora_ncsense_EH_true(p, C, S, w, yt, (float*)q);
    for (int64_t k = 0; k < Nt; ++k) r[k] = b[k] - (q[k] + (float)lambda * x[k]);
  } else {
    memset(x, 0, sizeof(float complex) * (size_t)Nt);
    memcpy(r, b, sizeof(float complex) * (size_t)Nt);
  }
  memcpy(pp, r, sizeof(float complex) * (size_t)Nt);
  double rn = 0;
  for (int64_t k = 0; k < Nt; ++k) rn += (double)crealf(r[k]) * crealf(r[k]) + (double)cimagf(r[k]) * cimagf(r[k]);
  int it = 0;
  for (;;) {
    double rel = bn > 0 ? sqrt(rn / bn) : 0.0;
    if (hist) hist[it] = rel;
    if (isnan(rel) || rel <= tol || it >= maxit) break;
    ora_ncsense_E(p, C, S, (const float*)pp, yt);
    ora_ncsense_EH_true(p, C, S, w, yt, (float*)q);
    double pq = 0;
    for (int64_t k = 0; k < Nt; ++k) {
      q[k] += (float)lambda * pp[k];
      pq += (double)crealf(pp[k]) * crealf(q[k]) + (double)cimagf(pp[k]) * cimagf(q[k]);
    }
    double alpha = rn / pq;
    double rn_new = 0;
    for (int64_t k = 0; k < Nt; ++k) {
      x[k] += (float)alpha * pp[k];
      r[k] -= (float)alpha * q[k];
      rn_new += (double)crealf(r[k]) * crealf(r[k]) + (double)cimagf(r[k]) * cimagf(r[k]);
    }
    double beta = rn_new / rn;
    rn = rn_new;
    for (int64_t k = 0; k < Nt; ++k) pp[k] = r[k] + (float)beta * pp[k];
    ++it;
    tmark();
  }
  free(b); free(pp); free(r); free(q); free(yt);
  return it;
}
