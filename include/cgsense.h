/*
 * cgsense.h -- C ABI of the MI355X (gfx950) non-Cartesian CG-SENSE library
 * (libcgsense_hip.so).
 *
 * One plan holds one trajectory (optionally split into frames of uneven sample
 * counts), its density weights, the coil sensitivities and every device buffer
 * the operators need.  All entry points return a status code (CGS_OK == 0) and
 * never throw.  A plan is not re-entrant (the same rule as codeare's operators,
 * whose const methods write mutable scratch, NCSENSE.hpp:485-487); distinct
 * plans are independent.  Each plan enqueues its work on one HIP stream.
 *
 * Layouts are codeare's (column-major Matrix, src/matrix/Matrix.hpp:579):
 *   image   x[i0 + N0*(i1 + N1*i2)]                      complex64 (re,im)
 *   frames  x[i + Ntot*f]                                (N^d, F)
 *   data    y[j + M_f*c] inside frame f's block, blocks at C*frame_offsets[f]
 *           (one frame: the (M, C) matrix of NCSENSE::Trafo, NCSENSE.hpp:173)
 *   maps    S[i + Ntot*c]                                (N^d, C)
 *   k-space k[t + d*j], k(t, j) in [-0.5, 0.5) pairs with image axis d-1-t
 *           (NFFT.hpp:267-268 feeds NFFT3's row-major plan.x)
 *
 * By default every data pointer given to an operator is host memory (codeare
 * Matrix storage) and the call returns when the result is in it.  With
 * CGS_DEVICE_IO the pointers are device memory, calls are asynchronous on the
 * plan's stream, and cgs_sync() waits.
 */
#ifndef CGSENSE_H
#define CGSENSE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cgs_plan_s* cgs_plan;

/* status codes (codeare::error_code analogue, src/common.h:30-60) */
enum {
  CGS_OK = 0,
  CGS_ERR_ARG = 1,          /* bad argument (shape, size, null pointer) */
  CGS_ERR_STATE = 2,        /* call out of order (e.g. no trajectory yet) */
  CGS_ERR_HIP = 3,          /* HIP runtime error */
  CGS_ERR_ALLOC = 4,        /* device allocation failed */
  CGS_ERR_UNSUPPORTED = 5,  /* parameters outside what the kernels implement */
  CGS_ERR_NAN = 6           /* NaN in a solver scalar (CGLS.hpp:41 break) */
};

/* plan flags */
#define CGS_DEVICE_IO      0x1u  /* operator data pointers are device memory */
#define CGS_PROFILE        0x2u  /* time every kernel launch with HIP events */
#define CGS_DETERMINISTIC  0x4u  /* fixed summation order in the gridding (slower) */
#define CGS_AUTOTUNE       0x8u  /* cgs_set_trajectory runs (or reuses) cgs_autotune */
#define CGS_MAPS_PER_FRAME 0x10u /* one set of maps per frame: cgs_set_sensitivities
                                    takes (N^d, F, C), the (N0, N1, Nz, C) maps of
                                    codeare's 3rd_dim_cart slices (NCSENSE.hpp:98-99) */

/* E^H variants */
#define CGS_EH_TRUE        0     /* x = sum_c conj(S_c) .* NFFT^H(w .* y_c)   */
#define CGS_EH_REF         1     /* x = I .* sum_c conj(S_c) .* ITrafo_c(y_c) (NCSENSE.hpp:289-298) */

/* CG operator E^H W E of the true-adjoint solve */
#define CGS_NORMAL_NUFFT      0  /* E then E^H: window gather, FFTs, gridding scatter */
#define CGS_NORMAL_TOEPLITZ   1  /* Toeplitz embedding: pad, FFT_2N, .* kernel, IFFT_2N, crop */

/* stopping rules of cgs_solve */
#define CGS_TOL_NORM_RATIO    0  /* ||r|| / ||E^H y|| <= tol                   */
#define CGS_TOL_SQUARED_RATIO 1  /* rn / xn <= tol, codeare's cgeps (CGLS.hpp:41) */

typedef struct {
  int maxit;          /* CG iterations at most (codeare cgiter / cgmaxit)   */
  double tol;         /* stopping threshold, see tol_kind (codeare cgeps)   */
  int tol_kind;       /* CGS_TOL_NORM_RATIO or CGS_TOL_SQUARED_RATIO        */
  double lambda;      /* Tikhonov weight (codeare lambda, CGLS.hpp:48-49)   */
  int eh_mode;        /* CGS_EH_TRUE or CGS_EH_REF                          */
  int ftiter;         /* CGNR steps of the reference E^H (codeare ftiter)   */
  double fteps;       /* CGNR absolute break on the weighted residual       */
  int normal_op;      /* CGS_NORMAL_NUFFT or CGS_NORMAL_TOEPLITZ (true E^H) */
  int graph_iters;    /* >0: capture this many CG iterations (true E^H) in one
                         HIP graph and replay it; 0: launch kernels one by one */
  int joint;          /* 1: one CG over all frames together -- alpha, beta and the
                         stop test from sums over every frame, as codeare's CGLS
                         runs on the whole (N^d, F) matrix for nmany frames and
                         3rd_dim_cart slices (CGLS.hpp:30-59 on NCSENSE.hpp:251-348);
                         0: every frame has its own scalars and stops on its own */
} cgs_solve_opts;

typedef struct {
  int iters[1024];      /* iterations run, per frame (first 1024 frames)       */
  double final_rel[1024];/* final residual measure per frame (tol_kind units)   */
  int nframes;
  int max_iters;
  int status;           /* CGS_OK, or CGS_ERR_NAN if a frame hit a NaN scalar */
  double done_ms[1024]; /* per frame: host-observed time from the call's start
                           to the convergence check that saw the frame stop
                           (one launch step late at most; for latency figures) */
  double total_ms;      /* whole call, upload and download included          */
} cgs_solve_report;

/* --- plan life cycle (NFFT.hpp:66-191, NCSENSE.hpp:69-184) --------------- */

/* d = 2 or 3; N[d] image sides (codeare imsz); M_total samples over all frames;
 * sigma = oversampling (codeare alpha, n = ceil(sigma*N)); m = KB cut-off;
 * ncoils = C; nframes = F with frame_offsets[F+1] sample offsets (NULL when
 * F == 1).  Replaces constructing C NFFT<cxfl> plans plus NCSENSE<cxfl>. */
int cgs_plan_create(cgs_plan* plan, int d, const int* N, int64_t M_total,
                    double sigma, int m, int ncoils, int nframes,
                    const int64_t* frame_offsets, unsigned flags);
int cgs_plan_destroy(cgs_plan plan);

/* Use this stream (hipStream_t) instead of the plan's own. */
int cgs_set_stream(cgs_plan plan, void* hip_stream);
int cgs_sync(cgs_plan plan);

/* NCSENSE::KSpace / NFFT::KSpace (NCSENSE.hpp:204-236, NFFT.hpp:258-273).
 * k: host (d, M_total) float.  Sorts samples into grid tiles and precomputes
 * the window (the PRE_PSI analogue of NFFT::Weights, NFFT.hpp:281-296). */
int cgs_set_trajectory(cgs_plan plan, const float* k);
/* NCSENSE::Weights / NFFT::Weights (NCSENSE.hpp:244-248, NFFT.hpp:281-296):
 * host (M_total) float density weights; also fixes the damping mask
 * (NFFTTraits.hpp:225-260). */
int cgs_set_weights(cgs_plan plan, const float* w);
/* NCSENSE::Sensitivities (NCSENSE.hpp:193-197): host (N^d, C) complex64 maps,
 * or (N^d, F, C) with CGS_MAPS_PER_FRAME; also computes IntensityMap
 * (MRI.hpp:6-30), per frame with per-frame maps. */
int cgs_set_sensitivities(cgs_plan plan, const void* S);
/* The same with S a device pointer (same layout), loaded and the intensity map
 * computed on the plan's stream, bit-identical to cgs_set_sensitivities; the
 * caller keeps S alive until the stream has passed the call.  No counterpart
 * in the reference: for callers that keep many map sets on the device and
 * switch between them per solve (training loops, batched series). */
int cgs_set_sensitivities_device(cgs_plan plan, const void* S);

/* Off-resonance model of the NFFT operators: NFFT<cxfl> with "b0" and "timing"
 * (NFFT.hpp:117-173, NFFTTraits.hpp:91-108,150-166,188-216, NFFT3's
 * mri_inh_3d plan).  b0: host (N0, N1) float field (cycles per unit of t), t:
 * host (M_total) float sample times.  cgs_nfft_forward / _adjoint /
 * _inverse_ref then apply
 *   f_j = sum_r x(r) exp(-2 pi i k_j.r) exp(-2 pi i b0(r) t_j)
 * (and its adjoint, and the CGNR on it), with the time-frequency term as a
 * third NFFT dimension of oversampling 1.2 and window cut-off m.  2D, one
 * frame; needs the trajectory first (a later cgs_set_trajectory rebuilds the
 * model).  b0 = NULL removes it.  The NCSENSE operators (cgs_E, cgs_EH,
 * cgs_normal, cgs_solve) refuse a plan with b0 set: codeare's NCSENSE reads b0
 * but gives it to none of its NFFTs (NCSENSE.hpp:100-103). */
int cgs_set_b0(cgs_plan plan, const float* b0, const float* t);
/* Third-axis length N3, frequency scale omega and mid time ts of the model. */
int cgs_b0_info(cgs_plan plan, int* N3, double* omega, double* ts);

/* --- NFFT<cxfl> operators (one frame at a time, batch images) -------------- */
/* nfft_trafo of `batch` images (N^d, batch) -> (M_f, batch) for frame `frame`
 * (NFFT::Trafo, NFFT.hpp:305-361).  batch <= ncoils. */
int cgs_nfft_forward(cgs_plan plan, int frame, const void* img, void* f, int batch);
/* true adjoint nfft_adjoint (NFFTTraits.hpp:201-204), no weights. */
int cgs_nfft_adjoint(cgs_plan plan, int frame, const void* f, void* img, int batch);
/* NFFT::Adjoint semantics: `maxit` steps of the weighted, damped CGNR
 * (NFFT.hpp:370-436, NFFTTraits.hpp:122-138), per image. */
int cgs_nfft_inverse_ref(cgs_plan plan, int frame, const void* f, void* img,
                         int batch, int maxit, double eps);

/* --- NCSENSE<cxfl> operators (all frames) ----------------------------------- */
/* E (NCSENSE::Trafo / operator*, NCSENSE.hpp:310-348, :423-425) */
int cgs_E(cgs_plan plan, const void* x, void* y);
/* E^H: CGS_EH_TRUE, or CGS_EH_REF = NCSENSE::operator/ (NCSENSE.hpp:251-301) */
int cgs_EH(cgs_plan plan, const void* y, void* x, int eh_mode, int ftiter, double fteps);
/* E^H W E x + lambda x with the true adjoint (one CG operator application),
 * through E and E^H or through the Toeplitz embedding. */
int cgs_normal(cgs_plan plan, const void* x, void* q, double lambda, int normal_op);

/* CG solve (NCSENSE::Adjoint / operator->* -> CGLS::Solve, NCSENSE.hpp:408-414,
 * CGLS.hpp:30-59).  x0 may be NULL (zero start, as codeare) or a warm start.
 * Every frame has its own scalars and stops on its own, or, with opts->joint,
 * all frames share one CG.  maxit = 0 returns E^H y as codeare's CGLS does
 * (CGLS.hpp:31-33), with or without x0.  report may be NULL. */
int cgs_solve(cgs_plan plan, const void* y, const void* x0, void* x,
              const cgs_solve_opts* opts, cgs_solve_report* report);

/* Allocation-free cgs_solve.  No counterpart in codeare: it is for callers that
 * cannot allocate device memory in the call (tight serving loops, code that
 * must not call hipMalloc).  Same arguments and results as cgs_solve on a
 * CGS_DEVICE_IO plan; nothing is allocated: the CG bookkeeping (per-frame
 * scalars and stop flags, reduction partials, the residual history) lives in a
 * fixed device array of CGS_NOALLOC_SCRATCH_BYTES per GPU, held until the
 * process ends, and every other buffer is the plan's.  Limits: graph_iters
 * must be 0 (instantiating a graph allocates); the Toeplitz operator needs its
 * spectrum prepared beforehand (one cgs_normal or cgs_solve with
 * CGS_NORMAL_TOEPLITZ on the plan); frames x (maxit + 1) history entries plus
 * the partials must fit the array (CGS_ERR_ARG otherwise; 8 MiB hold e.g. one
 * frame up to ~1e6 iterations, or 64 frames of 16 coils up to ~9000).
 * Every call on one GPU shares that array (it keeps the last call's
 * contents, and cgs_residual_history reads the last call's history from it):
 * calls that overlap on one GPU race. */
#define CGS_NOALLOC_SCRATCH_BYTES (8u << 20)
int cgs_solve_noalloc(cgs_plan plan, const void* y, const void* x0, void* x,
                      const cgs_solve_opts* opts, cgs_solve_report* report);

/* Residual measure (tol_kind units) of the last cgs_solve for one frame, one
 * entry per iteration starting with the initial one (CGLS's _res history,
 * CGLS.hpp:38, here per solve).  Returns the number of entries written. */
int cgs_residual_history(cgs_plan plan, int frame, double* out, int cap);

/* --- launch-shape autotuner (no counterpart in codeare) --------------------- */
typedef struct {
  int visits_per_item;   /* gridding: tile visits per work item (chunking)      */
  int grid_coils;        /* gridding: coils per register-accumulator pass       */
  int interp_coils;      /* gather: coils per LDS-staged pass                   */
  int from_cache;        /* 1: taken from an earlier plan of the same bucket     */
  double grid_ms, interp_ms;                  /* chosen configuration             */
  double default_grid_ms, default_interp_ms;  /* the plan's defaults              */
  int grid_mfma;         /* gridding: 1 matrix-core scatter, 0 VALU registers    */
  int grid_sweep;        /* gridding: 1 line sweep (2D), 0 the tile kernels      */
  int grid_bin;          /* gridding: 1 binned LDS scatter (each sample once)    */
  int interp_lanes;      /* gather lanes: 0 (coil, tap row); 2D 1, 2 (coil, sample), 3, 4 (coil pair, sample);
                            3D 1 (sample, tap plane), 2 (sample, tap plane, tap parity) */
} cgs_tune_report;
/* Times the gridding scatter (the binned LDS scatter; 2D: the line sweep; the tile
 * kernels -- matrix-core or VALU-register, visits per work item, coils per pass) and the
 * window gather of this plan's
 * trajectory over a small grid of launch shapes (`reps` timed runs each) and
 * keeps the fastest.  Results are cached per process under (d, n, C, frames, samples
 * per frame rounded to a power of two); with use_cache != 0 a cached bucket is
 * applied without timing.  Needs the trajectory; clobbers the plan's
 * intermediate buffers (not the trajectory, weights, maps or Toeplitz data). */
int cgs_autotune(cgs_plan plan, int reps, int use_cache, cgs_tune_report* report);

/* --- profiling (CGS_PROFILE plans) ------------------------------------------ */
/* kernel ids for cgs_kernel_stats */
enum {
  CGS_K_FFT_FWD = 0,   /* deapodise/pad/coil-multiply + forward FFT passes */
  CGS_K_INTERP = 1,    /* window gather (forward B)                        */
  CGS_K_GRID = 2,      /* tile-binned window scatter (adjoint B^T)          */
  CGS_K_FFT_ADJ = 3,   /* adjoint FFT passes + crop/deapodise               */
  CGS_K_COIL = 4,      /* conj(S) coil combine                              */
  CGS_K_CG = 5,        /* CG / CGNR vector and scalar kernels              */
  CGS_K_OTHER = 6,     /* layout conversions                               */
  CGS_K_TOEP_PAD = 7,  /* Toeplitz: x.*S_c, zero-pad to 2N, FFT along axis 0  */
  CGS_K_TOEP_CONV = 8, /* Toeplitz: FFT, .* kernel spectrum, IFFT along the
                          last axis (one fused pass)                        */
  CGS_K_TOEP_CROP = 9, /* Toeplitz: IFFT along axis 0 (and 1 in 3D), crop    */
  CGS_K_COUNT = 10
};
/* Accumulated launches, total milliseconds and algorithmic bytes per kernel
 * class since the last reset (reset = 1 clears after reading). */
int cgs_kernel_stats(cgs_plan plan, int kernel, int64_t* launches, double* ms,
                     double* bytes, int reset);
/* Time only the kernel classes whose bit (1 << id) is set (default: all);
 * the others are still counted.  Fewer events, less perturbation. */
int cgs_profile_mask(cgs_plan plan, unsigned mask);
/* Sizes the plan resolved: n[3] oversampled sides, tile side, sorted sample
 * count, gridding work items. */
int cgs_plan_info(cgs_plan plan, int* n, int* tile, int64_t* visits, int* grid_items);
/* The gridding scatter the plan runs now (after CGS_GRID_* switches or
 * cgs_autotune): CGS_GRID_TILES (tile owners: matrix-core or VALU register
 * kernels), CGS_GRID_SWEEP_LINES (2D line sweep) or CGS_GRID_BINNED (each sample
 * once into an LDS region per bin).  No counterpart in the reference. */
enum { CGS_GRID_TILES = 0, CGS_GRID_SWEEP_LINES = 1, CGS_GRID_BINNED = 2 };
int cgs_gridding_kind(cgs_plan plan);
/* Average duration of the Toeplitz convolution pass (the CG iteration's
 * dominant kernel) over `reps` back-to-back launches between two events, on a
 * zeroed scratch copy of its buffer (the plan's state is untouched); needs a 2D
 * plan whose Toeplitz spectrum is prepared.  No counterpart in the reference:
 * a measurement aid for roofline figures. */
int cgs_time_toeplitz_conv(cgs_plan plan, int reps, double* us_per_launch);

const char* cgs_status_string(int status);
const char* cgs_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* CGSENSE_H */
