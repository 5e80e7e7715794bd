/*
 * aarambam.h -- C-ABI of the MI355X NG-2LPT initial-condition library
 * (libaarambam_hip.so).
 *
 * The library is a drop-in for the hot path of the reference binary
 * Aarambam/LPT/Basis (`2LPTBasis <ParameterFile>`): it keeps the parameter
 * file, the alpha / X0 coefficient tables, the transfer table, the glass
 * file, the seed-table semantics and the Gadget / potential-dump outputs,
 * and runs displacement_fields() on one GPU.  Plain C types only.
 *
 * Reference interfaces each entry point replaces (paths relative to the
 * reference repository root):
 *   aar_read_params      Aarambam/LPT/Basis/read_param.c:8-262 (read_parameterfile)
 *   aar_read_params_res  Aarambam/LPT/ResBasis/read_param.c:8-274 (read_parameterfile)
 *   aar_load_alpha       Aarambam/LPT/Basis/main.c:215-247 (AlphaTable load + symmetrisation)
 *   aar_load_x0          Aarambam/LPT/Basis/main.c:250-289 (X0Table load)
 *   aar_load_transfer    Aarambam/LPT/Basis/power.c:66-153 (read_transfer_table)
 *   aar_read_glass       Aarambam/LPT/Basis/read_glass.c:10-88, 226-260 (read_glass, find_files)
 *   aar_cosmology        Aarambam/LPT/Basis/main.c:1547-1552 (set_units),
 *                        power.c:391-448 (initialize_powerspectrum), main.c:305-313, 372
 *   aar_seedtable        Aarambam/LPT/Basis/main.c:315-349
 *   aar_job_prepare/run  Aarambam/LPT/Basis/main.c:149-1532 (displacement_fields)
 *   aar_job_reseed       Aarambam/LPT/Basis/main.c:317-349 (seed table of a new Seed)
 *   aar_write_gadget     Aarambam/LPT/Basis/save.c:108-520 (write_particle_data, save_local_data)
 *   aar_write_gadget_files  Aarambam/LPT/Basis/save.c:108-520 with several writer
 *                        tasks (FileBase.<i>, num_files), save.c:184-187, 252
 *   aar_write_phi        Aarambam/LPT/Basis/save.c:14-105 (write_phi)
 *   aar_task_layout, aar_write_gadget_tasks, aar_write_phi_tasks
 *                        the same writers as run on several MPI tasks (main.c:1556-1599,
 *                        read_glass.c:129-146, save.c:14-520)
 *   aar_write_inputspec  Aarambam/LPT/Basis/main.c:1631-1694 (print_spec)
 *
 * Status codes: 0 on success; values below 100000 are the reference's
 * FatalError numbers where one applies; see aar_status_string().
 */
#ifndef AARAMBAM_H
#define AARAMBAM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AAR_STR 500

/* Status codes.  Numbers below 100000 are the reference's FatalError codes
 * (Aarambam/LPT/Basis/main.c:1612-1618 and the call sites listed there). */
enum {
  AAR_OK = 0,
  AAR_ERR_ALLOC = 1,               /* main.c ASSERT_ALLOC -> FatalError(1) */
  AAR_ERR_TABLE_FILE = 17,         /* missing alpha / X0 / transfer file (main.c:230,270; power.c:81) */
  AAR_ERR_KLOWER = 1111,           /* transfer table does not reach k < 10^-4.6 (power.c:117-121) */
  AAR_ERR_WRITE_PARALLEL = 24131,  /* NumFilesWrittenInParallel > NTask (save.c:112-117) */
  AAR_ERR_BAD_ARG = 100001,        /* invalid argument at the C-ABI */
  AAR_ERR_HIP = 100002,            /* a HIP runtime call failed */
  AAR_ERR_HBM_BUDGET = 100003,     /* job does not fit the admitted HBM budget */
  AAR_ERR_NMAX = 100004,           /* Nmax < Nsample (main.c:200 assert) */
  AAR_ERR_UNSUPPORTED = 100005,    /* a configuration the reference does not run either */
  AAR_ERR_NO_DEVICE = 100006,
  AAR_ERR_SCRATCH = 100007         /* job exceeds the fixed scratch of aar_job_run_noalloc */
};

/* The 38 tags of the reference parameter file (read_param.c:27-184), and the
 * three resonance tags of the ResBasis variant (Aarambam/LPT/ResBasis/
 * read_param.c:88-98).  Variant 0 is Aarambam/LPT/Basis, variant 1
 * Aarambam/LPT/ResBasis: its P(k) carries the oscillation
 * 1 + A_res cos(w_res k + phi_res) (WhichSpectrum 1) or
 * 1 + A_res cos(w_res ln k + phi_res) (WhichSpectrum 2), ResBasis/power.c:240-266,
 * and its Gaussian potential the modulation chain of ResBasis/main.c:445-455. */
typedef struct aar_params {
  int Nmesh, Nsample, Nmax, N_modes, Seed;
  int SphereMode, FixedAmplitude, PhaseFlip, SavePotentialField;
  int WhichSpectrum, WhichTransfer, GlassTileFac, NumFilesWrittenInParallel;
  int WDM_On, WDM_Vtherm_On;
  double Box, Omega, OmegaLambda, OmegaBaryon, OmegaDM_2ndSpecies, HubbleParam, Sigma8;
  double PrimordialIndex, Redshift, Fnl, ShapeGamma;
  double UnitLength_in_cm, UnitMass_in_g, UnitVelocity_in_cm_per_s;
  double InputSpectrum_UnitLength_in_cm, WDM_PartMass_in_kev;
  char GlassFile[AAR_STR], FileWithInputSpectrum[AAR_STR], FileWithInputTransfer[AAR_STR];
  char FileWithAlphaCoeff[AAR_STR], FileWithX0Coeff[AAR_STR], OutputDir[AAR_STR], FileBase[AAR_STR];
  int Variant;                   /* 0 Basis, 1 ResBasis */
  double A_res, w_res, phi_res;  /* ResBasis only (0 for Basis) */
} aar_params;

/* Host-side inputs of one job.  Arrays are owned by the caller. */
typedef struct aar_tables {
  const double *alpha;      /* N_modes^3, symmetrised, alpha[(i*Nm + j)*Nm + k] */
  const double *x0;         /* N_modes x 2 (X0Table rows) */
  int ntransfer;            /* WhichTransfer 2: rows of the transfer table */
  const double *transfer_k; /* raw k column */
  const double *transfer_t; /* raw T column */
  const float *glass;       /* nglass x 3 glass positions, file order */
  int nglass;
  double glass_box;         /* BoxSize from the glass header */
  const uint8_t *glass_header; /* optional 256-byte glass header (fill bytes are kept) */
} aar_tables;

/* Derived scalars (set_units, initialize_powerspectrum, displacement_fields). */
typedef struct aar_cosmo {
  double UnitTime_in_s, G, Hubble;
  double InitTime, Norm, sigma8_in, D0, Dplus, Anorm;
  double Beta, hubble_a, vel_prefac, vel_prefac2, fac;
} aar_cosmo;

enum { AAR_FP64 = 64, AAR_FP32 = 32 };
/* PLAIN: pairwise (N_i, N_j) schedule of the reference, one product and one
 * R2C per pair.  BATCHED: the sum regrouped by outer index K, all mode
 * fields and all K products in batched FFT launches. */
enum { AAR_SCHED_PLAIN = 0, AAR_SCHED_BATCHED = 1 };

/* Fields that aar_job_fetch_field can copy back, in the reference's
 * layouts: spectra as N x N x (N/2+1) complex (2 doubles each), real
 * fields as N x N x N doubles (padding dropped). */
enum {
  AAR_FIELD_PHIG_SPEC = 0,   /* phi_G(k) after main.c:524 */
  AAR_FIELD_PHI_SPEC = 1,    /* Phi(k) after main.c:1150 */
  AAR_FIELD_DISP_X = 2,      /* ZA displacement fields, real (main.c:1380) */
  AAR_FIELD_DISP_Y = 3,
  AAR_FIELD_DISP_Z = 4,
  AAR_FIELD_DISP2_X = 5,     /* 2LPT displacement fields, real, not divided by N^3 (main.c:1381) */
  AAR_FIELD_DISP2_Y = 6,
  AAR_FIELD_DISP2_Z = 7,
  AAR_FIELD_PHIG_REAL = 8,   /* write_phi(pot, 0) field (main.c:1096) */
  AAR_FIELD_PHING_REAL = 9,  /* write_phi(pot, 1) field (main.c:1135) */
};

typedef struct aar_ctx aar_ctx;
typedef struct aar_job aar_job;

const char *aar_status_string(int status);

/* ---- host side (no GPU needed) ---- */
int aar_read_params(const char *path, aar_params *out, char *msg, size_t msglen);
/* The ResBasis tag set (41 tags); sets out->Variant = 1. */
int aar_read_params_res(const char *path, aar_params *out, char *msg, size_t msglen);
int aar_load_alpha(const char *path, int n_modes, double *alpha_out /* Nm^3 */);
int aar_load_x0(const char *path, int n_modes, double *x0_out /* Nm*2 */);
/* count first (k, t NULL), then fill */
int aar_load_transfer(const char *path, int *nrows, double *k, double *t);
/* count first (pos NULL), then fill; header_out (256 bytes) optional */
int aar_read_glass(const char *path, int *nglass, float *pos, double *box, uint8_t *header_out);
int aar_cosmology(const aar_params *p, const aar_tables *t, aar_cosmo *out);
int aar_seedtable(int nmesh, int seed, uint32_t *out /* nmesh^2 */);
/* The same table filled by up to `threads` host threads (aar_seedtable uses up to eight from
 * nmesh 512 on): each jumps the one ranlxd stream to the first draw of its share (the
 * subtract-with-borrow recursion as a congruential generator modulo 2^576 - 2^240 + 1), so the
 * table is bit-identical for every thread count.  No counterpart in the reference, whose
 * main.c:317-349 draws the table serially. */
int aar_seedtable_threads(int nmesh, int seed, uint32_t *out, int threads);
/* n-th output of gsl_rng_get for ranlxd (luxury 202 or 397), GSL rng_test semantics */
unsigned long aar_ranlxd_nth(unsigned long seed, int luxury, int n);
int aar_write_gadget(const aar_params *p, const aar_cosmo *c, const float *pos, const float *vel,
                     long long npart, const uint8_t *glass_header);
/* The snapshot in nfiles files OutputDir/FileBase.<i> (one file: OutputDir/FileBase), the
 * layout the reference writes with several tasks: npart per file, npartTotal and num_files of
 * the run, IDs counting on across files.  nfiles 0 picks the fewest files whose records stay
 * within max_record bytes (0: 2^31 - 1, the limit of Gadget-1's 4-byte record markers, which a
 * single 1024^3 file overflows).  *nfiles_out (optional) receives the count. */
int aar_write_gadget_files(const aar_params *p, const aar_cosmo *c, const float *pos, const float *vel,
                           long long npart, const uint8_t *glass_header, int nfiles, long long max_record,
                           int *nfiles_out);
int aar_write_phi(const aar_params *p, const double *real_nnn, int non_gaussian);
/* The outputs as the reference writes them when run on ntask MPI tasks
 * (x-slabs of FFTW 2's MPI decomposition: blocks of ceil(Nmesh / ntask) planes,
 * main.c:1556-1599; particles to the task of their unperturbed x,
 * read_glass.c:129-146).  aar_task_layout: particles per task and the number of
 * tasks holding any (NTaskWithN).  aar_write_gadget_tasks: OutputDir/FileBase.<task>
 * for every task with particles, num_files = NTaskWithN, run-wide IDs
 * (save.c:108-520); FatalError 24131 when NumFilesWrittenInParallel > ntask.
 * aar_write_phi_tasks: one dump per task, ..._StartX_%04d_NumX_%d when
 * NTaskWithN > 1 (save.c:14-105). */
int aar_task_layout(const aar_params *p, const aar_tables *t, int ntask, long long *npart_task /* ntask, or NULL */,
                    int *ntask_with_n);
int aar_write_gadget_tasks(const aar_params *p, const aar_cosmo *c, const aar_tables *t, const float *pos,
                           const float *vel, long long npart, int ntask);
int aar_write_phi_tasks(const aar_params *p, const double *real_nnn, int non_gaussian, int ntask, int ntask_with_n);
int aar_write_inputspec(const aar_params *p, const aar_tables *t, const aar_cosmo *c);

/* ---- device side ---- */
int aar_ctx_create(int device, aar_ctx **out);
int aar_ctx_destroy(aar_ctx *ctx);
/* wait for all work on the context's device (hipDeviceSynchronize) */
int aar_ctx_sync(aar_ctx *ctx);
/* HBM bytes a job of this shape allocates (admission control): its own buffers (field
 * slots, seed table, particles) plus its tables (k-space tables, modal schedule, transfer
 * table, glass), which the context keeps resident and shares with later jobs of the same
 * parameters (Seed and Fnl aside), tables, precision and schedule. */
long long aar_job_bytes(const aar_params *p, const aar_tables *t, int precision, int schedule);
/* Number and device bytes of the job tables the context holds for later jobs (at most 32;
 * the least recently used one that no job holds is dropped first).  No counterpart in the
 * reference, which reads its tables once per process (main.c:215-289, power.c:66-153,
 * read_glass.c:10-88). */
int aar_ctx_cached_plans(aar_ctx *ctx, int *count, long long *bytes);
int aar_job_prepare(aar_ctx *ctx, const aar_params *p, const aar_tables *t, int precision,
                    int schedule, aar_job **out);
/* enqueue the whole job on `stream` (a hipStream_t, NULL = the context stream) */
int aar_job_run(aar_job *job, void *stream);
/* Make the next aar_job_run produce the IC of another Seed and Fnl with the same tables and
 * buffers, as a fresh aar_job_prepare would (main.c:317-349 builds the seed table from Seed;
 * Fnl enters at main.c:1104-1126).  The seed table is built on the host at this call, so it
 * overlaps with earlier runs still executing on the device; it is copied to the device at the
 * start of the next run, ahead of that run's kernels.  Results of the previous run must be
 * fetched before that next run is enqueued. */
int aar_job_reseed(aar_job *job, int seed, double fnl);
/* One whole IC job that allocates nothing: no hipMalloc, hipFree or other allocating device
 * routine.  No counterpart in the reference; it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid it).  It
 * computes what aar_job_prepare + aar_job_run compute for the same inputs, bit for bit, with
 * every device buffer besides the inputs and outputs (tables, seed table, field slots) taken
 * from a 64 MiB scratch array the library holds on each GPU until the process ends (enough for
 * Nmesh <= 64 in either precision and schedule); larger jobs get AAR_ERR_SCRATCH.  pos and vel
 * are the caller's device arrays of 3 npart floats (npart = GlassTileFac^3 nglass).  The work
 * is enqueued on `stream` (NULL: the context stream) and the call returns when it has finished.
 * Every call on one GPU shares that GPU's scratch, whose contents persist from the last call:
 * calls that overlap on one GPU (other threads, other contexts of the device) race. */
int aar_job_run_noalloc(aar_ctx *ctx, const aar_params *p, const aar_tables *t, int precision, int schedule,
                        float *pos, float *vel, void *stream);
/* wait for the job's last run (returns at once if it never ran) */
int aar_job_sync(aar_job *job);
/* device pointers of the result arrays (npart*3 floats each), valid until destroy */
int aar_job_device_output(aar_job *job, float **pos, float **vel, long long *npart);
int aar_job_fetch(aar_job *job, float *pos, float *vel);
int aar_job_fetch_field(aar_job *job, int which, double *out);
/* A sub-box of a real field (which >= AAR_FIELD_DISP_X): out[(x*ny + y)*nz + z]
 * for x in [x0, x0+nx) etc., without copying the whole grid back. */
int aar_job_fetch_field_box(aar_job *job, int which, int x0, int y0, int z0, int nx, int ny, int nz,
                            double *out);
/* seconds spent in the last run (HIP events around the enqueued work) */
double aar_job_last_seconds(aar_job *job);
/* Per-kernel-kind timing: when on, every launch of the next runs is bracketed
 * by HIP events on the job's stream; stats give launches, summed seconds and
 * summed algorithmic bytes (reads + writes of the data each launch must move). */
/* AAR_K_XY_SPLIT: the batched schedule's multi-output x / y pass pairs of the 2LPT spectra
 * (aar_xy_split_group_cols), whose algorithmic bytes count each input read once; AAR_K_MODES_XY:
 * the mode fields filtered in their x pass's loads (aar_ctx_set_modes_x) and their y passes,
 * phi_G read once. */
enum {
  AAR_K_GAUSS = 0, AAR_K_XY = 1, AAR_K_Z = 2, AAR_K_MODES = 3, AAR_K_PRODUCTS = 4,
  AAR_K_COMBINE = 5, AAR_K_LPT = 6, AAR_K_CIC = 7, AAR_K_ACCUM = 8, AAR_K_XY_SPLIT = 9, AAR_K_MODES_XY = 10,
  AAR_NKINDS = 11
};
int aar_job_set_profiling(aar_job *job, int on);
/* Graph mode: the first run captures every launch of the job into one
 * hipGraph, later runs replay it.  Per-kind profiling applies to eager runs only. */
int aar_job_set_graph(aar_job *job, int on);
int aar_job_kernel_stats(aar_job *job, int kind, int *launches, double *seconds, double *bytes);
int aar_job_destroy(aar_job *job);
/* The batched schedule's fused z pass of the NG sum (z C2R of the mode fields, products, z R2C)
 * is compiled at aar_job_prepare for the job's own product schedule (hiprtc; cached per device
 * and schedule, so one compile serves every job of a table), with the same operations in the
 * same order as the library's kernel: results are bit-identical either way.  No counterpart in
 * the reference (FFTW's planner is the nearest).  aar_ctx_set_jit(ctx, 0) keeps later jobs of
 * the context on the library's kernel.  aar_job_jit: 1 if the job runs a compiled kernel, 0 if
 * it runs the library's (pairwise schedule, other mesh sizes, the allocation-free path, or off),
 * else the status of a compile that failed (the library's kernel then runs). */
int aar_ctx_set_jit(aar_ctx *ctx, int on);
int aar_job_jit(const aar_job *job);
/* Batched jobs at powers of two N >= 64 can filter their mode fields (q_n(k) phi_G(k),
 * main.c:571-639) in the loads of the x pass that takes them towards real space, so the mode
 * fields are never stored in k-space; each value's filter is then evaluated on its own, where
 * the separate filter pass advances one Legendre recurrence through all the fields.  mode 1 (the
 * default): when the recurrences are short (at most two steps per field on average, as for the
 * local and equilateral tables); 2: always; 0: never.  The same operations in the same order
 * either way: results are bit-identical.  No counterpart in the reference. */
int aar_ctx_set_modes_x(aar_ctx *ctx, int mode);

/* ---- template check of a finished job (no counterpart in the reference) ----
 * Binned statistics of the potentials on the device.  Bins are |k| in
 * [edges[b], edges[b+1]) on the job's k grid (2 pi / Box units), excluding the
 * origin, the Nyquist planes and the modes outside the Nsample cut.
 *   pk_meas[b]:   mean |phi_G(k)|^2 over the bin (all k, both halves);
 *   pk_theory[b]: mean of its expectation P(k) = Anorm k^ns (fac Beta / k^2)^2;
 *   kmean[b], nmodes[b]: mean |k| and number of modes in the bin;
 *   b_meas[(b1 nb + b2) nb + b3]: N^3 sum over closed triangles (k1 in b1, ...)
 *     of phi_G(k1) phi_G(k2) psi(k3), psi = (Phi - phi_G) N^3 / Fnl the
 *     non-Gaussian part built from the modal table;
 *   b_theory: its expectation from P(k) and the table (tree level);
 *   ntri: number of closed triangles of the bin triple.
 * Any output pointer may be NULL.  nb <= 32. */
int aar_job_ng_statistics(aar_job *job, int nb, const double *edges, double *pk_meas, double *pk_theory,
                          double *kmean, double *nmodes, double *b_meas, double *b_theory, double *ntri);

/* ---- modal decomposition on the device ----
 * The two dense contractions of the decomposition that produces the alpha
 * tables, with the reference's argument lists (host arrays):
 *   aar_mode_bispectra_covariance  Aarambam/basis/tetrapyd.c:111-208
 *     (compute_mode_bispectra_covariance): cov[n*n_modes + m] =
 *     sum_t w_t Q_n(t) Q_m(t), Q_n the symmetrised product of the 1-D mode
 *     functions mode_evals[p*k_npts + i] at the tetrapyd point
 *     (tetra_i1, i2, i3)[t]; runs on the FP64 matrix cores.
 *   aar_mode_QS                    Aarambam/basis/tetrapyd.c:210-281 (compute_QS):
 *     QS[s*n_modes + n] = sum_t S[s*tetra_npts + t] w_t Q_n(t).
 * Indices outside [0, k_npts) or [0, mode_p_max) are rejected
 * (AAR_ERR_BAD_ARG).  Limits from the LDS the kernels use: mode_p_max <= 39 for
 * the covariance, mode_p_max * k_npts <= 7936 for QS (AAR_ERR_UNSUPPORTED). */
int aar_mode_bispectra_covariance(aar_ctx *ctx, double *mode_bispectra_covariance, const double *tetra_weights,
                                  const int *tetra_i1, const int *tetra_i2, const int *tetra_i3, int tetra_npts,
                                  const int *mode_p1, const int *mode_p2, const int *mode_p3, int n_modes,
                                  const double *mode_evals, int mode_p_max, int k_npts);
int aar_mode_QS(aar_ctx *ctx, double *QS, const double *S, int n_shapes, const double *tetra_weights,
                const int *tetra_i1, const int *tetra_i2, const int *tetra_i3, int tetra_npts, const int *mode_p1,
                const int *mode_p2, const int *mode_p3, int n_modes, const double *mode_evals, int mode_p_max,
                int k_npts);
/* Basis.basis_expansion (Aarambam/basis/cmbbest.pyx:320-490): the expansion
 * of n_shapes shapes S[s*tetra_npts + t] in the normalised mode bispectra by
 * orthogonal matching pursuit.  The modes with p1 <= 3, p2 <= 3, p3 <= 3 start
 * selected (all with max_modes -1); each step solves the normal equations of
 * the selection by conjugate gradients (scipy's cg: x0 = 0, stop when
 * |r| < rtol |b|, at most niter iterations, plus regularization on the
 * diagonal), and adds, of the ten unselected modes most correlated with the
 * residual, the one that raises the condition number of the selection's Gram
 * matrix least; it stops when |R|^2 / <S,S> <= rr_tol, when max_modes - 1
 * modes are selected (max_modes 0: n_modes), or when the condition number
 * exceeds cond_tol; a final solve with rtol 1e-16 (10^6 iterations) gives
 * the coefficients.  Modes (p,0,0) with p <= 2, (1,1,0) and those flagged in
 * bad_modes (n_modes bytes, or NULL) are dropped.  As in the reference, the
 * residual that drives the selection is that of the first shape for every
 * shape.  Outputs: alpha[s*n_modes + n], shape_cov (n_shapes^2, may be NULL),
 * corr and mse per shape (the reference's convergence_correlation and
 * convergence_MSE, may be NULL), selected[s*n_modes + n] (may be NULL).  QQ,
 * QS, the residuals and their projections run on the device, and so do the
 * condition numbers (np.linalg.cond in the reference): the extreme eigenvalue
 * magnitudes of the selection's Gram matrix and of the ten bordered ones, one
 * block each (Householder tridiagonalisation, then bisection on Sturm
 * counts); the conjugate gradients run on the host. */
int aar_mode_basis_expansion(aar_ctx *ctx, double *alpha, double *shape_cov, double *corr, double *mse,
                             uint8_t *selected, const double *S, int n_shapes, const double *tetra_weights,
                             const int *tetra_i1, const int *tetra_i2, const int *tetra_i3, int tetra_npts,
                             const int *mode_p1, const int *mode_p2, const int *mode_p3, int n_modes,
                             const double *mode_evals, int mode_p_max, int k_npts, const uint8_t *bad_modes,
                             double regularization, long long niter, double rtol, double rr_tol, double cond_tol,
                             int max_modes);

/* ---- serving: continuous batching of IC jobs on one device ----
 * No counterpart in the reference, which runs one IC per process
 * (main.c:53-102); every job keeps the semantics of aar_job_prepare/run.
 * A server allocates one HBM arena of hbm_budget bytes once and runs up to
 * max_active (1..16) jobs at a time, each on its own stream, carved from the
 * arena by footprint (aar_job_bytes).  Jobs are admitted in submission order
 * as slots and arena space free up, so jobs submitted while others run
 * overlap with them on the device.  Results are bit-identical to the same
 * job run alone. */
typedef struct aar_server aar_server;
int aar_server_create(aar_ctx *ctx, long long hbm_budget, int max_active, aar_server **out);
/* Queue a job (tables are copied; the job's own buffers come from the arena, its tables from
 * the context's cache of device-resident tables, outside the arena).  keep = 0: its arena piece is returned as
 * soon as it completes and results are dropped; keep = 1: results stay on the
 * device until aar_server_release.  AAR_ERR_HBM_BUDGET if the job can never fit. */
int aar_server_submit(aar_server *s, const aar_params *p, const aar_tables *t, int precision, int schedule,
                      int keep, long long *ticket);
/* Block until the job has completed; latency_s = completion - submission,
 * queue_s = admission - submission.  Returns the job's status. */
int aar_server_wait(aar_server *s, long long ticket, double *latency_s, double *queue_s);
int aar_server_fetch(aar_server *s, long long ticket, float *pos, float *vel);
int aar_server_device_output(aar_server *s, long long ticket, float **pos, float **vel, long long *npart);
/* Forget a completed job and return a kept job's arena piece. */
int aar_server_release(aar_server *s, long long ticket);
int aar_server_stats(aar_server *s, long long *submitted, long long *completed, int *peak_active,
                     long long *peak_bytes);
/* Cancels jobs not yet admitted, waits for running ones, frees the arena. */
int aar_server_destroy(aar_server *s);

/* ---- autotuner of the FFT tile shapes ----
 * Counterpart of the reference's FFT planning (rfftw3d_mpi_create_plan,
 * main.c:1556-1599): times each tile shape of the x and y passes on nf fields
 * of an n^3 grid and keeps the fastest per axis for later jobs.  Every shape
 * uses the same radix plan, so the choice never changes results.  A job's
 * first aar_job_prepare at n >= 256 tunes by itself on its own buffers.
 * Then it times the radix plans of the x / y passes at n = 512 and 1024 (aar_autotune_plans) and
 * the column groups of the x / y pass pairs (whole passes over all fields, or
 * field by field over all columns, halves, quarters) and keeps the fastest (aar_xy_group_cols).
 * seconds: optional [2][4] mean seconds per launch (-1: shape not built for n);
 * variant: optional [2] chosen shape per axis. */
int aar_autotune(aar_ctx *ctx, int n, int precision, int nf, double *seconds, int *variant);
/* aar_autotune, also reporting the radix plans of the x / y passes it timed at n = 512 and 1024:
 * plan_seconds (optional, 3 entries, -1 where not timed) mean seconds of one x + y pass over nf
 * fields per plan; plan (optional) the plan kept.  Plans: 0 radices spread evenly (512 = 8 8 8,
 * 1024 = 16 8 8), 1 radix 16 first (16 16 2, 16 16 4), 2 radix 16 last (2 16 16, 4 16 16).  Unlike
 * the tile shapes, plans change the rounding of the transforms (by a few units in the last place).
 * Counterpart of FFTW's planner choosing among algorithms (rfftw3d_mpi_create_plan, main.c:1566). */
int aar_autotune_plans(aar_ctx *ctx, int n, int precision, int nf, double *plan_seconds, int *plan);
/* aar_autotune reporting every timing of its one search (each optional, -1 where not timed):
 * tile_seconds [2][4] as aar_autotune's; pair_seconds [3][4] mean seconds of one x + y pass over
 * nf fields per radix plan and column group (whole passes, all columns, halves, quarters);
 * split_seconds [9] mean seconds per multi-output pair call over whole passes, all columns,
 * halves, quarters, eighths, then at the group kept per tile shape 0-3 of the pairs' y pass.  The
 * choices kept: aar_xy_tile_variant (axes 0, 1, 2), aar_xy_radix_plan, aar_xy_group_cols,
 * aar_xy_split_group_cols. */
int aar_autotune_report(aar_ctx *ctx, int n, int precision, int nf, double *tile_seconds,
                        double *pair_seconds, double *split_seconds);
/* radix plan in use for (n, precision), or a negative status; the setter pins one (0 only for
 * meshes other than 512 and 1024) */
int aar_xy_radix_plan(int n, int precision);
int aar_set_xy_radix_plan(int n, int precision, int plan);
/* tile shape in use for (n, precision, axis 0 = x, 1 = y, 2 = the y pass of the multi-output
 * pairs, whose loads also apply the output factors) */
int aar_xy_tile_variant(int n, int precision, int axis);
/* pin a tile shape (0, 1, 2, 3; for axis 2 also -1: follow axis 1); a shape not built for n falls
 * back to 0 at launch.  Every shape gives bit-identical results. */
int aar_set_xy_tile_variant(int n, int precision, int axis, int variant);
/* Column groups of the x / y pass pairs, also picked by aar_autotune: 0 runs each pass over all
 * fields at once; cols > 0 (a multiple of 16, at most n / 2) runs them field by field over kz
 * columns [k0, k0 + cols), so a group's intermediate can stay in the last-level cache between
 * the two passes.  Results are bit-identical for every choice.  The getter returns the choice
 * in use for (n, precision) or a negative status. */
int aar_xy_group_cols(int n, int precision);
int aar_set_xy_group_cols(int n, int precision, int cols);
/* Column groups of the batched schedule's multi-output x / y pass pairs: the nine 2LPT spectra
 * (main.c:1158-1277, 1311-1369) are one spectrum times factors of kx, ky and kz, so they leave
 * k-space as five x transforms, each read by the y pass once per spectrum it feeds, times that
 * spectrum's ky and kz factors (the three 2LPT displacements as two x transforms).  aar_autotune
 * (and a job's first prepare) also times these pairs over whole passes, all columns, halves,
 * quarters and eighths on nine fields and keeps the fastest.  cols = -1 in the setter follows
 * the pairs' choice (aar_xy_group_cols).  Results are bit-identical for every choice.
 * Counterpart of FFTW's planning (main.c:1556-1599). */
int aar_xy_split_group_cols(int n, int precision);
int aar_set_xy_split_group_cols(int n, int precision, int cols);

/* ---- building blocks (tests, tools) ---- */
/* 3-D transforms with FFTW 2 rfftwnd conventions on a host array in the
 * reference layout: dir=+1 C2R (in: N*N*(N/2+1) complex, out: N^3 real),
 * dir=-1 R2C (in: N^3 real, out: complex).  precision AAR_FP64/AAR_FP32. */
int aar_fft3d(aar_ctx *ctx, int n, int precision, int dir, const double *in, double *out);
/* Gaussian potential phi_G(k) only (reference complex layout, doubles). */
int aar_gaussian_field(aar_ctx *ctx, const aar_params *p, const aar_tables *t, int precision,
                       double *out);
/* Mean seconds per launch of one FFT pass (kind 0 x, 1 y, 2 z-C2R, 3 z-R2C)
 * or of a plain streaming copy kernel (kind 4, reads and writes nf spectra) over nf fields of an N^3 grid;
 * kinds 5 / 6 read and write back the kz < N/2 block of the nf fields with the x / y pass's tile
 * addressing and no transform (the access pattern's own rate, a bound for the passes). */
int aar_bench_pass(aar_ctx *ctx, int n, int precision, int kind, int nf, int iters, double *sec);
/* Device ranlxd draws for tests: `n` uniforms of stream `seed`. */
int aar_ranlxd_device(aar_ctx *ctx, uint32_t seed, int n, double *out);
/* The condition-number kernel of aar_mode_basis_expansion on its own, for tests:
 * ext[2b] = min |lambda| and ext[2b + 1] = max |lambda| of the symmetric m x m
 * matrices mats[b*m*m ...] (host arrays, b < nmat).  m <= 1982. */
int aar_sym_eig_extremes(aar_ctx *ctx, int m, int nmat, const double *mats, double *ext);

#ifdef __cplusplus
}
#endif
#endif
