/*
 * cgs_oracle.h -- CPU restatement (FP64) of codeare's non-Cartesian CG-SENSE path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or the timed CPU port -- never as the product path.
 *
 * What it restates (citations are relative to the codeare source tree):
 *   NFFT<cxfl>::Trafo / nfft_trafo      src/matrix/ft/NFFT.hpp:305-361, NFFTTraits.hpp:176-179
 *   nfft_adjoint                        src/matrix/ft/NFFTTraits.hpp:201-204
 *   NFFT<cxfl>::Adjoint = ITrafo CGNR   src/matrix/ft/NFFT.hpp:370-436, NFFTTraits.hpp:122-138
 *   damping mask w_hat                  src/matrix/ft/NFFTTraits.hpp:225-260
 *   KB window (phi, phi_hut)            src/matrix/ft/NFFTTraits.hpp:659-673 (b = pi(2-1/sigma))
 *   IntensityMap                        src/matrix/mri/MRI.hpp:6-30
 *   NCSENSE E / reference E^H           src/matrix/ft/NCSENSE.hpp:251-348
 *   CGLS::Solve                         src/optimisation/CGLS.hpp:30-59
 *   dotc                                src/matrix/linalg/Lapack.hpp:619-624
 *
 * NFFT3 itself is not in the reference tree (third-party libnfft3, version
 * unpinned by the reference's CMake).  The NFFT here follows the published
 * NFFT3 method (Keiner, Kunis, Potts, ACM TOMS 36(4), 2009): Kaiser-Bessel
 * window, 2m+2 taps per dimension, deconvolution by 1/I0(.), FFT of size
 * n = ceil(sigma*N).  Its scaling is pinned only by the direct NDFT.
 *
 * Layout conventions (codeare Matrix is column-major, Matrix.hpp:579):
 *   image  img[i0 + N0*(i1 + N1*i2)], complex as interleaved (re,im)
 *   trajectory k[t + d*j], k(t,j) pairs with image axis d-1-t
 *   data   y[j + M*c]  (sample fastest, coil slowest)
 *   maps   S[i + Ntot*c]
 */
#ifndef CGS_ORACLE_H
#define CGS_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct ora_nfft ora_nfft;

/* ---- primitives ---------------------------------------------------------- */
double ora_bessel_i0(double x);
/* KB window in grid-cell units: phi(t) for a tap at distance t cells. */
double ora_kb_phi(double t, int m, double b);
/* KB Fourier coefficient I0(m*sqrt(b^2-(2*pi*k/n)^2)). */
double ora_kb_phihut(double k, int n, int m, double b);
/* In-place 1D complex FFT of length n (interleaved), sign -1 or +1, unnormalised. */
int ora_fft1d(double* x, int n, int sign);

/* ---- NFFT plan ------------------------------------------------------------ */
ora_nfft* ora_nfft_create(int d, const int* N, int64_t M, double sigma, int m);
void ora_nfft_destroy(ora_nfft* p);
int ora_nfft_n(const ora_nfft* p, int axis);
/* k: (d, M) float, codeare layout.  Precomputes the window (PRE_PSI analogue). */
void ora_nfft_set_trajectory(ora_nfft* p, const float* k);
void ora_nfft_trafo(const ora_nfft* p, const double* img, double* f);
void ora_nfft_adjoint(const ora_nfft* p, const double* f, double* img);

/* direct NDFT / adjoint NDFT with the same axis pairing (FP64, O(M*N^d)) */
void ora_ndft_trafo(int d, const int* N, int64_t M, const float* k,
                    const double* img, double* f);
void ora_ndft_adjoint(int d, const int* N, int64_t M, const float* k,
                      const double* f, double* img);

/* ---- NFFT3 solver semantics ------------------------------------------------ */
/* w_hat disk/ball mask (1 inside radius N0/2 about N0/2, else 0). */
void ora_damping_mask(int d, const int* N, double* what);
/* Weighted, damped CGNR of NFFTTraits::ITrafo; returns iterations run.
 * dot_r_hist (nullable) receives dot_r before each loop test (maxit+1). */
int ora_itrafo(const ora_nfft* p, const double* y, const double* w,
               const double* what, int maxit, double eps, double* fhat,
               double* dot_r_hist);

/* ---- NCSENSE (complex64 I/O like Matrix<cxfl>, FP64 inside) ---------------- */
void ora_intensity_map(int64_t nr, int C, const float* S, float* Imap);
void ora_ncsense_E(const ora_nfft* p, int C, const float* S, const float* x, float* y);
void ora_ncsense_EH_true(const ora_nfft* p, int C, const float* S, const float* w,
                         const float* y, float* x);
void ora_ncsense_EH_ref(const ora_nfft* p, int C, const float* S, const float* w,
                        int ftiter, double fteps, const float* y, float* x);

/* CGLS with codeare's semantics: complex64 vectors, serial complex64 dotc,
 * squared-ratio stop rn/xn <= cgeps.  mode 0 = true adjoint E^H W, 1 = reference
 * E^H (ITrafo per coil).  Returns iterations run; res_hist gets rn/xn (maxit+1). */
int ora_cgls_ref(const ora_nfft* p, int C, const float* S, const float* w,
                 int mode, int ftiter, double fteps, int maxit, double cgeps,
                 double lambda, const float* y, float* x, double* res_hist);

/* CG on (E^H W E + lambda) x = E^H W y with FP64 scalars and complex64 vectors,
 * optional warm start x0, stop when ||r||/||E^H W y|| <= tol (norm ratio).
 * Returns iterations run; res_hist (nullable) gets the norm ratio (maxit+1). */
/* Seconds since the start of the last ora_cg_true / ora_cgls_ref call: entry 0
 * after E^H y, entry i after iteration i.  Returns the number written. */
int ora_time_marks(double* out, int cap);
int ora_cg_true(const ora_nfft* p, int C, const float* S, const float* w,
                int maxit, double tol, double lambda, const float* y,
                const float* x0, float* x, double* res_hist);

/* The two solvers on F frames or Cartesian slices solved together (one CG on
 * the (N^d, F) image matrix, codeare's nmany and 3rd_dim_cart, NCSENSE.hpp:
 * 81-99, 141-151, 251-348): frame f uses plans[f], maps S + 2*f*Sfs floats
 * (Sfs = 0: shared maps), weights w + moff[f] and the (M_f, C) data block at
 * y + 2*C*moff[f]; x is (N^d, F). */
int ora_cgls_ref_frames(const ora_nfft* const* plans, int F, int C, const float* S, int64_t Sfs,
                        const float* w, const int64_t* moff, int mode, int ftiter, double fteps,
                        int maxit, double cgeps, double lambda, const float* y, float* x, double* res_hist);
int ora_cg_true_frames(const ora_nfft* const* plans, int F, int C, const float* S, int64_t Sfs,
                       const float* w, const int64_t* moff, int maxit, double tol, double lambda,
                       const float* y, float* x, double* res_hist);

/* threads used by the OpenMP regions (0 = runtime default) */
void ora_set_threads(int n);
int ora_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
