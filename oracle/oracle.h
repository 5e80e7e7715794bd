/*
 * CPU oracle for the NG-2LPT initial-condition stage.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the CPU restatement that the HIP
 * product path is compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product library never links,
 * loads or calls anything here.
 *
 * It restates, in plain C99 (double precision, long double accumulators,
 * OpenMP over grid lines), the algorithm of the reference C binary
 * Aarambam/LPT/Basis (main.c, power.c, nrsrc/, read_glass.c) for one MPI
 * rank owning the whole grid.  Every function cites the reference lines it
 * follows.  Two third-party libraries of the reference are absent here and
 * are restated from their published algorithms:
 *   - GSL gsl_rng_ranlxd1 (M. Luescher's second-generation RANLUX, 48-bit
 *     double-precision variant, luxury p = 202).  Pinned by GSL's own
 *     known-answer value for the same code at p = 397 (rng/test.c).
 *   - GSL gsl_sf_legendre_Pl (upward Bonnet recurrence for l >= 3).
 *   - FFTW 2.1.5 rfftwnd_mpi: replaced by a radix-2 complex FFT with FFTW's
 *     real-transform conventions (unnormalised, R2C sign -, C2R sign +,
 *     complex passes over x and y before the half-complex pass over z).
 *
 * Grid layout is the reference's: complex index ((i*N + j)*(N/2+1) + k),
 * padded-real index ((i*N + j)*2*(N/2+1) + k)   (main.c:456, main.c:772).
 */
#ifndef AARAMBAM_ORACLE_H
#define AARAMBAM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  /* integer tags (read_param.c:8-184) */
  int Nmesh, Nsample, Nmax, N_modes, Seed;
  int SphereMode, FixedAmplitude, PhaseFlip;
  int WhichSpectrum, WhichTransfer, GlassTileFac;
  int WDM_On;
  /* float tags */
  double Box, Omega, OmegaLambda, OmegaBaryon, HubbleParam, Sigma8;
  double PrimordialIndex, Redshift, Fnl, ShapeGamma;
  double UnitLength_in_cm, UnitMass_in_g, UnitVelocity_in_cm_per_s;
  double InputSpectrum_UnitLength_in_cm, WDM_PartMass_in_kev;
  /* transfer table rows (WhichTransfer 2), raw k and T as read from file */
  int ntransfer;
  const double *transfer_k, *transfer_t;
  /* ResBasis variant (Aarambam/LPT/ResBasis): Variant 1 and its three tags
   * (ResBasis/read_param.c:88-98); Variant 0 is Basis */
  int Variant;
  double A_res, w_res, phi_res;
} orc_params;

typedef struct {
  double UnitTime_in_s, G, Hubble;
  double InitTime, Norm, sigma8_in, D0, Dplus, Anorm;
  double Beta, hubble_a, vel_prefac, vel_prefac2, fac;
} orc_cosmo;

typedef struct {
  double xdbl[12];
  double carry;
  int ir, jr, ir_old, pr;
} orc_rng;

/* ---- third-party restatements ---- */
void orc_ranlxd_set(orc_rng *r, unsigned long seed, int luxury);
double orc_ranlxd_uniform(orc_rng *r);
unsigned long orc_ranlxd_get(orc_rng *r);
/* n-th (1-based) output of gsl_rng_get after seeding, as GSL's rng_test does */
unsigned long orc_ranlxd_nth_get(unsigned long seed, int luxury, int n);
void orc_ranlxd_draws(unsigned long seed, int luxury, int n, double *out);
double orc_legendre_Pl(int l, double x);

/* ---- reference path ---- */
void orc_seedtable(int N, int Seed, uint32_t *out);
int orc_cosmology(const orc_params *p, orc_cosmo *c);
double orc_transfer(const orc_params *p, double k);
double orc_qromb_growth(const orc_params *p, double a);
double orc_sigma2_tophat(const orc_params *p, double norm, double R);

/* alpha rows -> symmetrised N_modes^3 table (main.c:215-241) */
void orc_alpha_symmetrise(int Nm, int nrows, const int *n1, const int *n2,
                          const int *n3, const double *a, double *table);

/* FFTW2 rfftwnd conventions on the padded in-place layout */
void orc_c2r(int N, double *data);
void orc_r2c(int N, double *data);

/* Gaussian potential phi_G(k) into cpot (complex, 2*N*N*(N/2+1) doubles) */
void orc_gaussian(const orc_params *p, const orc_cosmo *c, const uint32_t *seedtable,
                  double *cpot);

/* Full modal NG sum + combine (main.c:564-1151).  cpot in: phi_G(k);
 * out: final potential Phi(k) after the two round trips.  phig_real and
 * phing_real (optional, padded real) receive the two write_phi fields. */
int orc_ng_potential(const orc_params *p, const orc_cosmo *c, const double *alpha,
                     const double *x0, double *cpot, double *phig_real, double *phing_real);

/* ZA + 2LPT fields (main.c:1158-1381): disp[3], disp2[3] padded real fields */
void orc_lpt_fields(const orc_params *p, const orc_cosmo *c, const double *cpot,
                    double *disp0, double *disp1, double *disp2,
                    double *d2x, double *d2y, double *d2z);

/* Particle load from a glass (read_glass.c:176-214) */
void orc_glass_tile(const orc_params *p, const float *glass, int nglass, double glass_box,
                    float *pos, int *ids);

/* CIC readout + update (main.c:1440-1516) */
void orc_particles(const orc_params *p, const orc_cosmo *c, long long npart,
                   double *const disp[3], double *const disp2[3], float *pos, float *vel);

/* WDM thermal velocities the snapshot writer adds (save.c:350-356, power.c:624-707): kicks
 * vel (3 npart floats, particle order) in place, drawing from drand48 reset to its default
 * state; orc_drand48 is that generator (POSIX's 48-bit LCG). */
void orc_wdm_thermal(const orc_params *p, long long npart, float *vel);
void orc_drand48_reset(void);
double orc_drand48(void);

/* Whole displacement_fields() for one rank; optional outputs may be NULL. */
int orc_run(const orc_params *p, const double *alpha, const double *x0, const float *glass,
            int nglass, double glass_box, float *pos, float *vel, int *ids,
            double *phig_spec, double *phi_spec, double *phig_real, double *phing_real);

/* Modal decomposition contractions (Aarambam/basis/tetrapyd.c:111-208, 210-281):
 * cov[n*n_modes + m] = sum_t w_t Q_n(t) Q_m(t) and QS[s*n_modes + n] =
 * sum_t S[s*npts + t] w_t Q_n(t), Q_n the symmetrised product of three 1-D
 * mode functions, accumulated over t in order in double. */
void orc_mode_covariance(double *cov, const double *w, const int *i1, const int *i2, const int *i3, int npts,
                         const int *p1, const int *p2, const int *p3, int n_modes, const double *evals, int pmax,
                         int nk);
void orc_mode_qs(double *qs, const double *S, int n_shapes, const double *w, const int *i1, const int *i2,
                 const int *i3, int npts, const int *p1, const int *p2, const int *p3, int n_modes,
                 const double *evals, int pmax, int nk);

#ifdef __cplusplus
}
#endif
#endif
