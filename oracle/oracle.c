/*
 * CPU oracle for the NG-2LPT initial-condition stage -- TEST INFRASTRUCTURE.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this.  See oracle.h for the scope and the third-party restatements.
 * Reference paths below are relative to the reference repository root.
 */
#include "oracle.h"

#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.14159265358979323846 /* Aarambam/LPT/Basis/allvars.h:4 */
#define ORC_GRAVITY 6.672e-8          /* allvars.h:5 */
#define ORC_HUBBLE 3.2407789e-18      /* allvars.h:6 */

/* ------------------------------------------------------------------------ */
/* GSL gsl_rng_ranlxd1 / ranlxd2 (GSL rng/ranlxd.c, after M. Luescher's      */
/* ranlxd v2).  Subtract-with-borrow x_n = x_{n-5} - x_{n-12} - c (mod 1) on */
/* 48-bit fractions held exactly in doubles; after every 12 outputs the      */
/* generator advances `pr` steps (202 for luxury level 1, 397 for level 2)   */
/* and returns the last 12 values in generation order.                       */
/* ------------------------------------------------------------------------ */
static const double ORC_ONE_BIT = 1.0 / 281474976710656.0; /* 2^-48 */

void orc_ranlxd_set(orc_rng *r, unsigned long s, int luxury) {
  int xbit[31];
  if (s == 0) s = 1; /* GSL maps seed 0 to 1 */
  long i = (long)(s & 0xFFFFFFFFUL);
  for (int k = 0; k < 31; ++k) {
    xbit[k] = (int)(i % 2);
    i /= 2;
  }
  int ibit = 0, jbit = 18;
  for (int k = 0; k < 12; ++k) {
    double x = 0;
    for (int l = 1; l <= 48; ++l) {
      double y = (double)((xbit[ibit] + 1) % 2);
      x += x + y;
      xbit[ibit] = (xbit[ibit] + xbit[jbit]) % 2;
      ibit = (ibit + 1) % 31;
      jbit = (jbit + 1) % 31;
    }
    r->xdbl[k] = ORC_ONE_BIT * x;
  }
  r->carry = 0;
  r->ir = 11;
  r->jr = 7;
  r->ir_old = 0;
  r->pr = luxury;
}

static void orc_ranlxd_increment(orc_rng *r) {
  int ir = r->ir, jr = r->jr;
  double carry = r->carry;
  for (int k = 0; k < r->pr; ++k) {
    double y = r->xdbl[jr] - r->xdbl[ir] - carry;
    if (y < 0) {
      carry = ORC_ONE_BIT;
      y += 1;
    } else {
      carry = 0;
    }
    r->xdbl[ir] = y;
    ir = (ir + 1) % 12;
    jr = (jr + 1) % 12;
  }
  r->ir = ir;
  r->ir_old = ir;
  r->jr = jr;
  r->carry = carry;
}

double orc_ranlxd_uniform(orc_rng *r) {
  r->ir = (r->ir + 1) % 12;
  if (r->ir == r->ir_old) orc_ranlxd_increment(r);
  return r->xdbl[r->ir];
}

unsigned long orc_ranlxd_get(orc_rng *r) {
  return (unsigned long)(orc_ranlxd_uniform(r) * 4294967296.0);
}

unsigned long orc_ranlxd_nth_get(unsigned long seed, int luxury, int n) {
  orc_rng r;
  unsigned long k = 0;
  orc_ranlxd_set(&r, seed, luxury);
  for (int i = 0; i < n; ++i) k = orc_ranlxd_get(&r);
  return k;
}

void orc_ranlxd_draws(unsigned long seed, int luxury, int n, double *out) {
  orc_rng r;
  orc_ranlxd_set(&r, seed, luxury);
  for (int i = 0; i < n; ++i) out[i] = orc_ranlxd_uniform(&r);
}

/* GSL gsl_sf_legendre_Pl (specfunc/legendre_poly.c): closed forms for l<=2,
 * exact values at x = +-1, upward recurrence otherwise. */
double orc_legendre_Pl(int l, double x) {
  if (l == 0) return 1.0;
  if (l == 1) return x;
  if (l == 2) return 0.5 * (3.0 * x * x - 1.0);
  if (x == 1.0) return 1.0;
  if (x == -1.0) return (l & 1) ? -1.0 : 1.0;
  double pm2 = 1.0, pm1 = x, p = x;
  for (int ell = 2; ell <= l; ++ell) {
    p = (x * (2 * ell - 1) * pm1 - (ell - 1) * pm2) / ell;
    pm2 = pm1;
    pm1 = p;
  }
  return p;
}

/* ------------------------------------------------------------------------ */
/* Seed table, main.c:315-349                                                 */
/* ------------------------------------------------------------------------ */
void orc_seedtable(int N, int Seed, uint32_t *st) {
  orc_rng r;
  orc_ranlxd_set(&r, (unsigned long)Seed, 202);
#define DRAW ((uint32_t)(0x7fffffff * orc_ranlxd_uniform(&r)))
  for (int i = 0; i < N / 2; i++) {
    for (int j = 0; j < i; j++) st[i * N + j] = DRAW;
    for (int j = 0; j < i + 1; j++) st[j * N + i] = DRAW;
    for (int j = 0; j < i; j++) st[(N - 1 - i) * N + j] = DRAW;
    for (int j = 0; j < i + 1; j++) st[(N - 1 - j) * N + i] = DRAW;
    for (int j = 0; j < i; j++) st[i * N + (N - 1 - j)] = DRAW;
    for (int j = 0; j < i + 1; j++) st[j * N + (N - 1 - i)] = DRAW;
    for (int j = 0; j < i; j++) st[(N - 1 - i) * N + (N - 1 - j)] = DRAW;
    for (int j = 0; j < i + 1; j++) st[(N - 1 - j) * N + (N - 1 - i)] = DRAW;
  }
#undef DRAW
}

/* ------------------------------------------------------------------------ */
/* Cosmology, power.c and nrsrc/                                             */
/* ------------------------------------------------------------------------ */
typedef struct {
  const orc_params *p;
  double Norm;
  double r_tophat;
  /* sorted tabulated transfer, normalised to T(k_min)=1 (power.c:66-153) */
  int nt;
  double *logk, *logT;
} orc_ctx;

static double tk_eh(const orc_params *p, double k) { /* power.c:519-548 */
  double hubble = p->HubbleParam, omegam = p->Omega;
  double ombh2 = p->OmegaBaryon * p->HubbleParam * p->HubbleParam;
  if (p->OmegaBaryon == 0) ombh2 = 0.04 * p->HubbleParam * p->HubbleParam;
  k *= (3.085678e24 / p->UnitLength_in_cm);
  double theta = 2.728 / 2.7;
  double ommh2 = omegam * hubble * hubble;
  double s = 44.5 * log(9.83 / ommh2) / sqrt(1. + 10. * exp(0.75 * log(ombh2))) * hubble;
  double a = 1. - 0.328 * log(431. * ommh2) * ombh2 / ommh2 +
             0.380 * log(22.3 * ommh2) * (ombh2 / ommh2) * (ombh2 / ommh2);
  double gamma = a + (1. - a) / (1. + exp(4 * log(0.43 * k * s)));
  gamma *= omegam * hubble;
  double q = k * theta * theta / gamma;
  double L0 = log(2. * exp(1.) + 1.8 * q);
  double C0 = 14.2 + 731. / (1. + 62.5 * q);
  return L0 / (L0 + C0 * q * q);
}

static int cmp_pair(const void *a, const void *b) {
  const double *x = (const double *)a, *y = (const double *)b;
  return (x[0] > y[0]) - (x[0] < y[0]);
}

static int ctx_init(orc_ctx *cx, const orc_params *p) {
  memset(cx, 0, sizeof(*cx));
  cx->p = p;
  if (p->WhichTransfer == 2) { /* power.c:66-153 */
    int n = p->ntransfer;
    if (n < 2) return 17;
    double *pairs = malloc(sizeof(double) * 2 * n);
    for (int i = 0; i < n; i++) {
      pairs[2 * i] = log10(p->transfer_k[i]);
      pairs[2 * i + 1] = log10(p->transfer_t[i]);
    }
    qsort(pairs, n, 2 * sizeof(double), cmp_pair);
    if (pairs[0] >= -4.6) {
      free(pairs);
      return 1111;
    }
    cx->nt = n;
    cx->logk = malloc(sizeof(double) * n);
    cx->logT = malloc(sizeof(double) * n);
    double tl = pairs[1];
    for (int i = 0; i < n; i++) {
      cx->logk[i] = pairs[2 * i];
      cx->logT[i] = pairs[2 * i + 1] - tl;
    }
    free(pairs);
  }
  return 0;
}

static void ctx_free(orc_ctx *cx) {
  free(cx->logk);
  free(cx->logT);
}

static double transfer_ctx(const orc_ctx *cx, double k) { /* power.c:32-63 */
  const orc_params *p = cx->p;
  switch (p->WhichTransfer) {
    case 1:
      return tk_eh(p, k);
    case 2: { /* power.c:181-217 */
      k *= (p->InputSpectrum_UnitLength_in_cm / p->UnitLength_in_cm);
      double logk = log10(k);
      if (logk < cx->logk[0] || logk > cx->logk[cx->nt - 1]) return 0;
      int lo = 0, hi = cx->nt - 1;
      while (hi - lo > 1) {
        int mid = (hi + lo) / 2;
        if (logk < cx->logk[mid])
          hi = mid;
        else
          lo = mid;
      }
      double dlogk = cx->logk[hi] - cx->logk[lo];
      double u = (logk - cx->logk[lo]) / dlogk;
      double logT = (1 - u) * cx->logT[lo] + u * cx->logT[hi];
      return pow(10.0, logT);
    }
    default:
      return 1.0; /* tk_Efstathiou, power.c:504-507 */
  }
}

double orc_transfer(const orc_params *p, double k) {
  orc_ctx cx;
  if (ctx_init(&cx, p)) return NAN;
  double t = transfer_ctx(&cx, k);
  ctx_free(&cx);
  return t;
}

/* WDM suppression of power.c:267-275 (Bode, Ostriker & Turok 2001, eq. A9, g_X = 1.5) */
static double wdm_factor(const orc_params *p, double k, double power) {
  if (p->WDM_On != 1) return power;
  double alpha = 0.048 * pow((p->Omega - p->OmegaBaryon) / 0.4, 0.15) * pow(p->HubbleParam / 0.65, 1.3) *
                 pow(1.0 / p->WDM_PartMass_in_kev, 1.15);
  double Tf = pow(1 + pow(alpha * k * (3.085678e24 / p->UnitLength_in_cm), 2 * 1.2), -5.0 / 1.2);
  return power * (Tf * Tf);
}

static double powerspec_base(const orc_ctx *cx, double k);
static double powerspec(const orc_ctx *cx, double k) { return wdm_factor(cx->p, k, powerspec_base(cx, k)); }

static double powerspec_base(const orc_ctx *cx, double k) { /* power.c:240-266 */
  const orc_params *p = cx->p;
  if (p->Variant == 1) { /* ResBasis/power.c:240-260 */
    double power = 0;
    switch (p->WhichSpectrum) {
      case 0: power = cx->Norm * exp(p->PrimordialIndex * log(k)); break;
      case 1:
        power = cx->Norm * exp(p->PrimordialIndex * log(k)) * (1 + p->A_res * cos(p->w_res * k + p->phi_res));
        break;
      case 2:
        power = cx->Norm * exp(p->PrimordialIndex * log(k)) *
                (1 + p->A_res * cos(p->w_res * log(k) + p->phi_res));
        break;
    }
    double t = transfer_ctx(cx, k);
    power *= t * t;
    return power;
  }
  if (p->WhichSpectrum == 0) {
    double t = transfer_ctx(cx, k);
    return cx->Norm * exp(p->PrimordialIndex * log(k)) * t * t;
  }
  /* WhichSpectrum 1: Eisenstein-Hu spectrum times k^(ns-1) */
  double t = tk_eh(p, k);
  return cx->Norm * k * pow(t, 2) * pow(k, p->PrimordialIndex - 1.0);
}

/* Numerical Recipes polint/trapzd/qromb (nrsrc/polint.c, trapzd.c, qromb.c),
 * 1-based arrays as in the book. */
static void nr_polint(const double xa[], const double ya[], int n, double x, double *y,
                      double *dy) {
  int ns = 1;
  double c[8], d[8];
  double dif = fabs(x - xa[1]);
  for (int i = 1; i <= n; i++) {
    double dift = fabs(x - xa[i]);
    if (dift < dif) {
      ns = i;
      dif = dift;
    }
    c[i] = ya[i];
    d[i] = ya[i];
  }
  *y = ya[ns--];
  for (int m = 1; m < n; m++) {
    for (int i = 1; i <= n - m; i++) {
      double ho = xa[i] - x, hp = xa[i + m] - x;
      double w = c[i + 1] - d[i];
      double den = w / (ho - hp);
      d[i] = hp * den;
      c[i] = ho * den;
    }
    *y += (*dy = (2 * ns < (n - m) ? c[ns + 1] : d[ns--]));
  }
}

typedef double (*orc_fn)(const orc_ctx *, double);

static double nr_trapzd(orc_fn f, const orc_ctx *cx, double a, double b, int n, double *s) {
  if (n == 1) return (*s = 0.5 * (b - a) * (f(cx, a) + f(cx, b)));
  int it = 1;
  for (int j = 1; j < n - 1; j++) it <<= 1;
  double tnm = it, del = (b - a) / tnm, x = a + 0.5 * del, sum = 0.0;
  for (int j = 1; j <= it; j++, x += del) sum += f(cx, x);
  *s = 0.5 * (*s + (b - a) * sum / tnm);
  return *s;
}

static double nr_qromb(orc_fn f, const orc_ctx *cx, double a, double b) {
  double s[42], h[43], ss = 0, dss = 0, st = 0;
  h[1] = 1.0;
  for (int j = 1; j <= 40; j++) {
    s[j] = nr_trapzd(f, cx, a, b, j, &st);
    if (j >= 5) {
      nr_polint(&h[j - 5], &s[j - 5], 5, 0.0, &ss, &dss);
      if (fabs(dss) <= 1.0e-8 * fabs(ss)) return ss;
    }
    h[j + 1] = 0.25 * h[j];
  }
  fprintf(stderr, "Too many steps in routine qromb\n");
  return 0.0;
}

static double sigma2_int(const orc_ctx *cx, double k) { /* power.c:561-576 */
  double kr = cx->r_tophat * k, kr2 = kr * kr, kr3 = kr2 * kr;
  if (kr < 1e-8) return 0;
  double w = 3 * (sin(kr) / kr3 - cos(kr) / kr2);
  return 4 * ORC_PI * k * k * w * w * powerspec(cx, k);
}

static double growth_int(const orc_ctx *cx, double a) { /* power.c:595-598 */
  const orc_params *p = cx->p;
  return pow(a / (p->Omega + (1 - p->Omega - p->OmegaLambda) * a + p->OmegaLambda * a * a * a),
             1.5);
}

static double growth(const orc_ctx *cx, double a) { /* power.c:585-592 */
  const orc_params *p = cx->p;
  double hubble_a = sqrt(p->Omega / (a * a * a) + (1 - p->Omega - p->OmegaLambda) / (a * a) +
                         p->OmegaLambda);
  return hubble_a * nr_qromb(growth_int, cx, 0, a);
}

double orc_qromb_growth(const orc_params *p, double a) {
  orc_ctx cx;
  ctx_init(&cx, p);
  double g = growth(&cx, a);
  ctx_free(&cx);
  return g;
}

double orc_sigma2_tophat(const orc_params *p, double norm, double R) {
  orc_ctx cx;
  ctx_init(&cx, p);
  cx.Norm = norm;
  cx.r_tophat = R;
  double s = nr_qromb(sigma2_int, &cx, 0, 500.0 * 1 / R);
  ctx_free(&cx);
  return s;
}

static double F_Omega(const orc_params *p, double a) { /* power.c:601-608 */
  double om = p->Omega / (p->Omega + a * (1 - p->Omega - p->OmegaLambda) +
                          a * a * a * p->OmegaLambda);
  return pow(om, 0.6);
}

static double F2_Omega(const orc_params *p, double a) { /* power.c:611-618 */
  double om = p->Omega / (p->Omega + a * (1 - p->Omega - p->OmegaLambda) +
                          a * a * a * p->OmegaLambda);
  return 2 * pow(om, 4. / 7.);
}

int orc_cosmology(const orc_params *p, orc_cosmo *c) {
  orc_ctx cx;
  int rc = ctx_init(&cx, p);
  if (rc) return rc;
  memset(c, 0, sizeof(*c));
  /* set_units, main.c:1547-1552 */
  c->UnitTime_in_s = p->UnitLength_in_cm / p->UnitVelocity_in_cm_per_s;
  c->G = ORC_GRAVITY / pow(p->UnitLength_in_cm, 3) * p->UnitMass_in_g * pow(c->UnitTime_in_s, 2);
  c->Hubble = ORC_HUBBLE * c->UnitTime_in_s;
  /* initialize_powerspectrum, power.c:391-448 */
  c->InitTime = 1 / (1 + p->Redshift);
  double R8 = 8 * (3.085678e24 / p->UnitLength_in_cm);
  cx.Norm = 1.0;
  cx.r_tophat = R8;
  double res = nr_qromb(sigma2_int, &cx, 0, 500.0 * 1 / R8);
  c->sigma8_in = sqrt(res);
  c->Norm = p->Sigma8 * p->Sigma8 / res;
  c->D0 = growth(&cx, 1.0) / growth(&cx, 1.e-8) * 1.e-8;
  c->Dplus = growth(&cx, 1.0) / growth(&cx, c->InitTime);
  /* Basis: power.c:426-446; ResBasis sets it for every spectrum, ResBasis/power.c:418 */
  c->Anorm = (p->Variant == 1 || p->WhichSpectrum == 0) ? c->Norm : 0.0;
  /* displacement_fields prefactors, main.c:305-313, 372 */
  double a = c->InitTime;
  c->hubble_a = c->Hubble * sqrt(p->Omega / pow(a, 3) + (1 - p->Omega - p->OmegaLambda) / pow(a, 2) +
                                 p->OmegaLambda);
  c->vel_prefac = a * c->hubble_a * F_Omega(p, a) / sqrt(a);
  c->vel_prefac2 = a * c->hubble_a * F2_Omega(p, a) / sqrt(a);
  c->fac = pow(2 * ORC_PI / p->Box, 1.5);
  c->Beta = 1.5 * p->Omega /
            (2998. * 2998. / p->UnitLength_in_cm / p->UnitLength_in_cm * 3.085678e24 * 3.085678e24) /
            c->D0;
  ctx_free(&cx);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* FFT with FFTW 2 rfftwnd conventions (unnormalised; C2R = complex inverse  */
/* over x and y, then half-complex -> real over z, imaginary parts of the z  */
/* bins 0 and N/2 ignored; R2C = the reverse with sign -1).                  */
/* ------------------------------------------------------------------------ */
typedef double complex cplx;

static cplx *twiddle_table(int n, int sign) {
  cplx *w = malloc(sizeof(cplx) * (n / 2 + 1));
  const long double tp = 6.283185307179586476925286766559005768L;
  for (int m = 0; m <= n / 2; m++) {
    long double ang = tp * (long double)m / (long double)n;
    w[m] = (double)cosl(ang) + I * (double)(sign * sinl(ang));
  }
  return w;
}

/* w has n/2 + 1 entries exp(sign*2*pi*i*m/n); the others by conjugate symmetry */
static inline cplx root_of(const cplx *w, int n, long e) { return e <= n / 2 ? w[e] : conj(w[n - e]); }

/* mixed-radix decimation in time for lengths that are not powers of two (FFTW takes any
   size): y = DFT_n of x[0], x[s], ..., x[(n-1)s], split by the smallest prime factor p of n
   into p transforms of n/p and combined with the roots of unity of the ntop-point table */
static void fft_mixed(const cplx *x, int n, int s, cplx *y, const cplx *w, int ntop) {
  if (n == 1) {
    y[0] = x[0];
    return;
  }
  int p = 2;
  while (n % p) p++;
  const int m = n / p;
  for (int r = 0; r < p; r++) fft_mixed(x + (size_t)r * s, m, s * p, y + (size_t)r * m, w, ntop);
  cplx *t = malloc(sizeof(cplx) * n);
  for (int k = 0; k < m; k++)
    for (int u = 0; u < p; u++) {
      cplx acc = 0;
      for (int r = 0; r < p; r++) acc += y[r * m + k] * root_of(w, ntop, ((long)r * (k + u * m) % n) * (ntop / n));
      t[k + u * m] = acc;
    }
  memcpy(y, t, sizeof(cplx) * n);
  free(t);
}

static void fft1d(cplx *a, int n, const cplx *w) {
  if (n & (n - 1)) {
    cplx *x = malloc(sizeof(cplx) * n);
    memcpy(x, a, sizeof(cplx) * n);
    fft_mixed(x, n, 1, a, w, n);
    free(x);
    return;
  }
  /* iterative radix-2, decimation in time; w[m] = exp(sign*2*pi*i*m/n) */
  for (int i = 1, j = 0; i < n; i++) {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      cplx t = a[i];
      a[i] = a[j];
      a[j] = t;
    }
  }
  for (int len = 2; len <= n; len <<= 1) {
    int step = n / len;
    for (int i = 0; i < n; i += len) {
      for (int k = 0; k < len / 2; k++) {
        cplx u = a[i + k], v = a[i + k + len / 2] * w[k * step];
        a[i + k] = u + v;
        a[i + k + len / 2] = u - v;
      }
    }
  }
}

static void complex_pass(int N, double *data, int axis, int sign) {
  const int H = N / 2 + 1;
  cplx *c = (cplx *)data;
  cplx *w = twiddle_table(N, sign);
#pragma omp parallel
  {
    cplx *buf = malloc(sizeof(cplx) * N);
#pragma omp for schedule(static)
    for (long line = 0; line < (long)N * H; line++) {
      int a = (int)(line / H), k = (int)(line % H);
      size_t base, stride;
      if (axis == 0) { /* along i, fixed (j=a, k) */
        base = (size_t)a * H + k;
        stride = (size_t)N * H;
      } else { /* along j, fixed (i=a, k) */
        base = (size_t)a * N * H + k;
        stride = H;
      }
      for (int t = 0; t < N; t++) buf[t] = c[base + t * stride];
      fft1d(buf, N, w);
      for (int t = 0; t < N; t++) c[base + t * stride] = buf[t];
    }
    free(buf);
  }
  free(w);
}

void orc_c2r(int N, double *data) {
  const int H = N / 2 + 1;
  complex_pass(N, data, 0, +1);
  complex_pass(N, data, 1, +1);
  cplx *w = twiddle_table(N, +1);
#pragma omp parallel
  {
    cplx *buf = malloc(sizeof(cplx) * N);
#pragma omp for schedule(static)
    for (long row = 0; row < (long)N * N; row++) {
      double *r = data + (size_t)row * 2 * H;
      buf[0] = r[0];
      buf[N / 2] = r[2 * (N / 2)];
      for (int m = 1; m < N / 2; m++) {
        buf[m] = r[2 * m] + I * r[2 * m + 1];
        buf[N - m] = conj(buf[m]);
      }
      fft1d(buf, N, w);
      for (int z = 0; z < N; z++) r[z] = creal(buf[z]);
      r[N] = 0;
      r[N + 1] = 0;
    }
    free(buf);
  }
  free(w);
}

void orc_r2c(int N, double *data) {
  const int H = N / 2 + 1;
  cplx *w = twiddle_table(N, -1);
#pragma omp parallel
  {
    cplx *buf = malloc(sizeof(cplx) * N);
#pragma omp for schedule(static)
    for (long row = 0; row < (long)N * N; row++) {
      double *r = data + (size_t)row * 2 * H;
      for (int z = 0; z < N; z++) buf[z] = r[z];
      fft1d(buf, N, w);
      for (int m = 0; m <= N / 2; m++) {
        r[2 * m] = creal(buf[m]);
        r[2 * m + 1] = cimag(buf[m]);
      }
    }
    free(buf);
  }
  free(w);
  complex_pass(N, data, 1, -1);
  complex_pass(N, data, 0, -1);
}

/* ------------------------------------------------------------------------ */
/* Shared k-space helpers (main.c:412-452, 578-606)                          */
/* ------------------------------------------------------------------------ */
static inline double kcomp(int N, int i, double Box) {
  return (i < N / 2) ? i * 2 * ORC_PI / Box : -(N - i) * 2 * ORC_PI / Box;
}

static inline int outside_cut(const orc_params *p, const double kv[3], double kmag1) {
  if (p->SphereMode == 1) return kmag1 * p->Box / (2 * ORC_PI) > p->Nsample / 2;
  if (fabs(kv[0]) * p->Box / (2 * ORC_PI) > p->Nsample / 2) return 1;
  if (fabs(kv[1]) * p->Box / (2 * ORC_PI) > p->Nsample / 2) return 1;
  if (fabs(kv[2]) * p->Box / (2 * ORC_PI) > p->Nsample / 2) return 1;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Gaussian potential, main.c:382-524 (one rank: Local_x_start=0, nx=N)      */
/* ------------------------------------------------------------------------ */
void orc_gaussian(const orc_params *p, const orc_cosmo *c, const uint32_t *seedtable,
                  double *cpot) {
  const int N = p->Nmesh, H = N / 2 + 1;
  const double phase_shift = (p->PhaseFlip == 1) ? ORC_PI : 0.0;
  memset(cpot, 0, sizeof(double) * 2 * (size_t)N * N * H);
#pragma omp parallel for schedule(dynamic, 4)
  for (int i = 0; i < N; i++) {
    orc_rng r;
    for (int j = 0; j < N; j++) {
      orc_ranlxd_set(&r, seedtable[i * N + j], 202);
      for (int k = 0; k < N / 2; k++) {
        double phase = orc_ranlxd_uniform(&r) * 2 * ORC_PI;
        phase += phase_shift;
        double ampl;
        do
          ampl = orc_ranlxd_uniform(&r);
        while (ampl == 0);
        if (i == N / 2 || j == N / 2 || k == N / 2) continue;
        if (i == 0 && j == 0 && k == 0) continue;
        double kv[3];
        kv[0] = kcomp(N, i, p->Box);
        kv[1] = kcomp(N, j, p->Box);
        kv[2] = kcomp(N, k, p->Box);
        double kmag2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
        double kmag = sqrt(kmag2);
        if (outside_cut(p, kv, kmag)) continue;
        double phig = c->Anorm * exp(p->PrimordialIndex * log(kmag));
        if (p->Variant == 1) switch (p->WhichSpectrum) { /* ResBasis/main.c:445-455, no breaks */
            case 0: phig *= 1.0; /* fall through */
            case 1: phig *= (1.0 + p->A_res * cos(p->w_res * kmag + p->phi_res)); /* fall through */
            case 2: phig *= (1.0 + p->A_res * cos(p->w_res * log(kmag) + p->phi_res));
          }
        if (!p->FixedAmplitude) phig *= -log(ampl);
        phig = sqrt(phig) * c->fac * c->Beta / kmag2;
        double re = phig * cos(phase), im = phig * sin(phase);
        if (k > 0) {
          size_t co = ((size_t)i * N + j) * H + k;
          cpot[2 * co] = re;
          cpot[2 * co + 1] = im;
        } else if (i == 0) {
          if (j >= N / 2) continue;
          int jj = N - j;
          size_t co = ((size_t)i * N + j) * H;
          cpot[2 * co] = re;
          cpot[2 * co + 1] = im;
          co = ((size_t)i * N + jj) * H;
          cpot[2 * co] = re;
          cpot[2 * co + 1] = -im;
        } else {
          if (i >= N / 2) continue;
          int ii = N - i, jj = N - j;
          if (ii == N) ii = 0;
          if (jj == N) jj = 0;
          size_t co = ((size_t)i * N + j) * H;
          cpot[2 * co] = re;
          cpot[2 * co + 1] = im;
          co = ((size_t)ii * N + jj) * H;
          cpot[2 * co] = re;
          cpot[2 * co + 1] = -im;
        }
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Modal non-Gaussian sum and combine, main.c:564-1151                       */
/* ------------------------------------------------------------------------ */
void orc_alpha_symmetrise(int Nm, int nrows, const int *n1, const int *n2, const int *n3,
                          const double *a, double *T) {
  memset(T, 0, sizeof(double) * Nm * Nm * Nm);
#define AT(x, y, z) T[((size_t)(x) * Nm + (y)) * Nm + (z)]
  for (int r = 0; r < nrows; r++) { /* main.c:231-241 */
    AT(n3[r], n2[r], n1[r]) = a[r];
    AT(n3[r], n1[r], n2[r]) = a[r];
    AT(n2[r], n3[r], n1[r]) = a[r];
    AT(n2[r], n1[r], n3[r]) = a[r];
    AT(n1[r], n3[r], n2[r]) = a[r];
    AT(n1[r], n2[r], n3[r]) = a[r];
  }
#undef AT
}

typedef struct {
  int N, H;
  size_t nc;
  double kmin_ns, kmax_ns, s;
} ng_geom;

/* per-mode radial quantities as the reference computes them in each loop */
static inline int mode_k(const orc_params *p, const ng_geom *g, int i, int j, int k, double *kmag1,
                         double *kmag2, double *xbar) {
  double kv[3];
  kv[0] = kcomp(g->N, i, p->Box);
  kv[1] = kcomp(g->N, j, p->Box);
  kv[2] = kcomp(g->N, k, p->Box);
  *kmag2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
  *kmag1 = sqrt(*kmag2);
  double kmag_ns = pow(*kmag1, (4 - p->PrimordialIndex) / 3);
  *xbar = -1 + 2 * (kmag_ns - g->kmin_ns) / (g->kmax_ns - g->kmin_ns);
  if (i + j + k == 0) return 0;
  return !outside_cut(p, kv, *kmag1);
}

/* Build mode field n into out (complex), following the N_i / N_j loops of
 * main.c:571-639 (and 666-731): n <= 3 power laws, n >= 4 Legendre fields
 * advanced one order per call by the field-level Bonnet recurrence held in
 * leg[0] (order n-3) and leg[1] (order n-2). */
static void build_mode_field(const orc_params *p, const ng_geom *g, const double *cpot,
                             const double *x0, int n, int seed_legendre, double *leg0,
                             double *leg1, double *out) {
  const int N = g->N, H = g->H;
  const double ns = p->PrimordialIndex;
  const double np = (double)n - 1;
  memset(out, 0, sizeof(double) * 2 * g->nc);
  double *next = NULL;
  if (n >= 4) next = calloc(2 * g->nc, sizeof(double));
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++)
      for (int k = 0; k <= N / 2; k++) {
        size_t co = ((size_t)i * N + j) * H + k;
        double kmag1, kmag2, xb;
        if (!mode_k(p, g, i, j, k, &kmag1, &kmag2, &xb)) continue;
        const double re = cpot[2 * co], im = cpot[2 * co + 1];
        if (n <= 3) {
          long double fnl_fac =
              pow(kmag1, (4 - ns) / 3 * ((double)n - 3.)) * pow(kmag1, 4 - ns);
          out[2 * co] = (double)(fnl_fac * re);
          out[2 * co + 1] = (double)(fnl_fac * im);
        } else {
          if (seed_legendre) {
            double Pl = orc_legendre_Pl(n - 1 - 2, xb) * pow(kmag1, (2 - ns));
            leg0[2 * co] = re * Pl;
            leg0[2 * co + 1] = im * Pl;
            Pl = orc_legendre_Pl(n - 1 - 1, xb) * pow(kmag1, (2 - ns));
            leg1[2 * co] = re * Pl;
            leg1[2 * co + 1] = im * Pl;
          }
          next[2 * co] = ((2 * np - 1) * xb * leg1[2 * co] - (np - 1) * leg0[2 * co]) / np;
          next[2 * co + 1] =
              ((2 * np - 1) * xb * leg1[2 * co + 1] - (np - 1) * leg0[2 * co + 1]) / np;
          double Pl = x0[2 * (n - 1)] * pow(kmag1, (2 - ns));
          out[2 * co] = next[2 * co] - re * Pl;
          out[2 * co + 1] = next[2 * co + 1] - im * Pl;
        }
      }
  if (n >= 4) { /* memcpy shifts, main.c:645-649 */
    memcpy(leg0, leg1, sizeof(double) * 2 * g->nc);
    memcpy(leg1, next, sizeof(double) * 2 * g->nc);
    free(next);
  }
}

static void to_padded_real(int N, const double *spec, double *padded) {
  /* the same buffer serves as complex and padded real in the reference */
  memcpy(padded, spec, sizeof(double) * 2 * (size_t)N * N * (N / 2 + 1));
}

int orc_ng_potential(const orc_params *p, const orc_cosmo *c, const double *alpha,
                     const double *x0, double *cpot, double *phig_real, double *phing_real) {
  const int N = p->Nmesh, H = N / 2 + 1, Nm = p->N_modes;
  if (p->Nmax < p->Nsample) return 200; /* main.c:200 assert */
  ng_geom g;
  g.N = N;
  g.H = H;
  g.nc = (size_t)N * N * H;
  double kmin = 2 * ORC_PI / p->Box, kmax = 2 * ORC_PI / p->Box * p->Nmax;
  g.kmin_ns = pow(kmin, (4 - p->PrimordialIndex) / 3);
  g.kmax_ns = pow(kmax, (4 - p->PrimordialIndex) / 3);
  const double ns = p->PrimordialIndex;
#define AL(a, b, cc) alpha[((size_t)(a) * Nm + (b)) * Nm + (cc)]
  double t_012 = -AL(1, 1, 1) / AL(0, 1, 2);
  if (AL(0, 1, 2) == 0) t_012 = 0;

  size_t nd = 2 * g.nc;
  long double *acc_re = calloc(g.nc, sizeof(long double));
  long double *acc_im = calloc(g.nc, sizeof(long double));
  double *fi = malloc(sizeof(double) * nd), *fj = malloc(sizeof(double) * nd);
  double *pt = malloc(sizeof(double) * nd);
  double *li0 = calloc(nd, sizeof(double)), *li1 = calloc(nd, sizeof(double));
  double *lj0 = calloc(nd, sizeof(double)), *lj1 = calloc(nd, sizeof(double));
  if (!acc_re || !acc_im || !fi || !fj || !pt || !li0 || !li1 || !lj0 || !lj1) return 1;

  for (int Ni = 0; Ni < Nm; Ni++) {
    build_mode_field(p, &g, cpot, x0, Ni, Ni == 4, li0, li1, fi);
    orc_c2r(N, fi);
    memset(lj0, 0, sizeof(double) * nd);
    memset(lj1, 0, sizeof(double) * nd);
    for (int Nj = Ni; Nj < Nm; Nj++) {
      build_mode_field(p, &g, cpot, x0, Nj, (Nj == 4) || (Ni >= 4 && Ni == Nj), lj0, lj1, fj);
      int ksum = 0; /* main.c:745-759 */
      for (int Nk = Nj; Nk < Nm; Nk++)
        if (fabs(AL(Ni, Nj, Nk)) > 0) ksum = 1;
      if (!ksum) continue;
      orc_c2r(N, fj);
      memset(pt, 0, sizeof(double) * nd);
#pragma omp parallel for schedule(static)
      for (long row = 0; row < (long)N * N; row++)
        for (int k = 0; k < N; k++) {
          size_t co = (size_t)row * 2 * H + k;
          pt[co] = fi[co] * fj[co];
        }
      orc_r2c(N, pt);
#pragma omp parallel for schedule(static)
      for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++)
          for (int k = 0; k <= N / 2; k++) { /* main.c:785-870 */
            size_t co = ((size_t)i * N + j) * H + k;
            double kmag1, kmag2, xb;
            if (!mode_k(p, &g, i, j, k, &kmag1, &kmag2, &xb)) continue;
            long double fnl_fac = 0;
            double a = NAN, b = NAN, cc, d;
            for (int Nk = Nj; Nk < Nm; Nk++) {
              double np = (double)Nk - 1;
              if (Nk <= 3) {
                fnl_fac += AL(Ni, Nj, Nk) * pow(kmag1, (4 - ns) / 3 * ((double)Nk - 3.));
              } else {
                if ((Nk == 4) || ((Nj >= 4) && (Nj == Nk))) {
                  a = orc_legendre_Pl(Nk - 1 - 2, xb) / kmag2;
                  b = orc_legendre_Pl(Nk - 1 - 1, xb) / kmag2;
                }
                cc = ((2 * np - 1) * xb * b - (np - 1) * a) / np;
                d = cc - x0[2 * (Nk - 1)] / kmag2;
                a = b;
                b = cc;
                fnl_fac += AL(Ni, Nj, Nk) * d;
              }
            }
            acc_re[co] += (long double)(fnl_fac * pt[2 * co]);
            acc_im[co] += (long double)(fnl_fac * pt[2 * co + 1]);
          }
    }
  }

  /* Beta divergence terms, main.c:884-1063 */
  for (int Ni = 0; Ni <= 3; Ni++) {
    build_mode_field(p, &g, cpot, x0, Ni, 0, li0, li1, fi);
    orc_c2r(N, fi);
    for (int Nj = Ni; Nj <= 3; Nj++) {
      build_mode_field(p, &g, cpot, x0, Nj, 0, lj0, lj1, fj);
      orc_c2r(N, fj);
      memset(pt, 0, sizeof(double) * nd);
#pragma omp parallel for schedule(static)
      for (long row = 0; row < (long)N * N; row++)
        for (int k = 0; k < N; k++) {
          size_t co = (size_t)row * 2 * H + k;
          pt[co] = fi[co] * fj[co];
        }
      orc_r2c(N, pt);
#pragma omp parallel for schedule(static)
      for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++)
          for (int k = 0; k <= N / 2; k++) {
            size_t co = ((size_t)i * N + j) * H + k;
            double kmag1, kmag2, xb;
            if (!mode_k(p, &g, i, j, k, &kmag1, &kmag2, &xb)) continue;
            long double fnl_fac = 0;
            if (Ni == 0 && Nj == 1)
              fnl_fac += -AL(0, 1, 2) * t_012 * pow(kmag1, (4 - ns) / 3 * (2 - 3.0));
            if (Ni == 0 && Nj == 2)
              fnl_fac += AL(0, 1, 2) * t_012 * pow(kmag1, (4 - ns) / 3 * (1 - 3.0));
            acc_re[co] += (long double)(fnl_fac * pt[2 * co]);
            acc_im[co] += (long double)(fnl_fac * pt[2 * co + 1]);
          }
    }
  }
#undef AL

  /* combine, main.c:1074-1151 */
  double *cng = fj;
  for (size_t co = 0; co < g.nc; co++) {
    cng[2 * co] = (double)acc_re[co];
    cng[2 * co + 1] = (double)acc_im[co];
  }
  orc_c2r(N, cpot);
  if (phig_real) to_padded_real(N, cpot, phig_real);
  orc_r2c(N, cpot);
  unsigned int nmesh3 = ((unsigned int)N) * ((unsigned int)N) * ((unsigned int)N);
  for (size_t co = 0; co < g.nc; co++) {
    cpot[2 * co] /= (double)nmesh3;
    cpot[2 * co + 1] /= (double)nmesh3;
    cpot[2 * co] += p->Fnl * cng[2 * co] / (double)nmesh3;
    cpot[2 * co + 1] += p->Fnl * cng[2 * co + 1] / (double)nmesh3;
  }
  cpot[0] = 0.;
  cpot[1] = 0.;
  orc_c2r(N, cpot);
  if (phing_real) to_padded_real(N, cpot, phing_real);
  orc_r2c(N, cpot);
  for (size_t co = 0; co < g.nc; co++) {
    cpot[2 * co] = cpot[2 * co] / (double)nmesh3;
    cpot[2 * co + 1] = cpot[2 * co + 1] / (double)nmesh3;
  }
  free(acc_re);
  free(acc_im);
  free(fi);
  free(fj);
  free(pt);
  free(li0);
  free(li1);
  free(lj0);
  free(lj1);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* ZA and 2LPT displacement fields, main.c:1158-1381                         */
/* ------------------------------------------------------------------------ */
void orc_lpt_fields(const orc_params *p, const orc_cosmo *c, const double *cpot, double *dx,
                    double *dy, double *dz, double *d2x, double *d2y, double *d2z) {
  const int N = p->Nmesh, H = N / 2 + 1;
  const size_t nc = (size_t)N * N * H, nd = 2 * nc;
  orc_ctx cx;
  ctx_init(&cx, p);
  double *cdisp[3] = {dx, dy, dz};
  double *cdisp2[3] = {d2x, d2y, d2z};
  double *dg[6];
  for (int a = 0; a < 6; a++) dg[a] = calloc(nd, sizeof(double));
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++)
      for (int k = 0; k <= N / 2; k++) { /* main.c:1170-1215 */
        size_t co = ((size_t)i * N + j) * H + k;
        double kv[3] = {kcomp(N, i, p->Box), kcomp(N, j, p->Box), kcomp(N, k, p->Box)};
        double kmag2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
        double kmag = sqrt(kmag2);
        double t_of_k = transfer_ctx(&cx, kmag);
        double twb = t_of_k / c->Dplus / c->Beta;
        for (int a = 0; a < 3; a++) {
          cdisp[a][2 * co + 1] = kv[a] * twb * cpot[2 * co];
          cdisp[a][2 * co] = -kv[a] * twb * cpot[2 * co + 1];
        }
      }
  static const int ga[6] = {0, 0, 0, 1, 1, 2}, gb[6] = {0, 1, 2, 1, 2, 2};
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++)
      for (int k = 0; k <= N / 2; k++) { /* main.c:1238-1277 */
        size_t co = ((size_t)i * N + j) * H + k;
        double kv[3] = {kcomp(N, i, p->Box), kcomp(N, j, p->Box), kcomp(N, k, p->Box)};
        for (int q = 0; q < 6; q++) {
          dg[q][2 * co] = -cdisp[ga[q]][2 * co + 1] * kv[gb[q]];
          dg[q][2 * co + 1] = cdisp[ga[q]][2 * co] * kv[gb[q]];
        }
      }
  for (int q = 0; q < 6; q++) orc_c2r(N, dg[q]);
#pragma omp parallel for schedule(static)
  for (long row = 0; row < (long)N * N; row++)
    for (int k = 0; k < N; k++) { /* main.c:1284-1294 */
      size_t co = (size_t)row * 2 * H + k;
      dg[3][co] = dg[0][co] * (dg[3][co] + dg[5][co]) + dg[3][co] * dg[5][co] -
                  dg[1][co] * dg[1][co] - dg[2][co] * dg[2][co] - dg[4][co] * dg[4][co];
    }
  orc_r2c(N, dg[3]);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++)
      for (int k = 0; k <= N / 2; k++) { /* main.c:1311-1369, CORRECT_CIC on */
        size_t co = ((size_t)i * N + j) * H + k;
        double kv[3] = {kcomp(N, i, p->Box), kcomp(N, j, p->Box), kcomp(N, k, p->Box)};
        double kmag2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
        double fx = 1, fy = 1, fz = 1;
        if (kv[0] != 0) {
          fx = (kv[0] * p->Box / 2) / N;
          fx = sin(fx) / fx;
        }
        if (kv[1] != 0) {
          fy = (kv[1] * p->Box / 2) / N;
          fy = sin(fy) / fy;
        }
        if (kv[2] != 0) {
          fz = (kv[2] * p->Box / 2) / N;
          fz = sin(fz) / fz;
        }
        double ff = 1 / (fx * fy * fz);
        double smth = ff * ff;
        for (int a = 0; a < 3; a++) {
          if (kmag2 > 0.0) {
            cdisp2[a][2 * co] = dg[3][2 * co + 1] * kv[a] / kmag2;
            cdisp2[a][2 * co + 1] = -dg[3][2 * co] * kv[a] / kmag2;
          } else {
            cdisp2[a][2 * co] = cdisp2[a][2 * co + 1] = 0.0;
          }
          cdisp[a][2 * co] *= smth;
          cdisp[a][2 * co + 1] *= smth;
          cdisp2[a][2 * co] *= smth;
          cdisp2[a][2 * co + 1] *= smth;
        }
      }
  for (int a = 0; a < 3; a++) {
    orc_c2r(N, cdisp[a]);
    orc_c2r(N, cdisp2[a]);
  }
  for (int a = 0; a < 6; a++) free(dg[a]);
  ctx_free(&cx);
}

/* ------------------------------------------------------------------------ */
/* Particle load, read_glass.c:176-214                                       */
/* ------------------------------------------------------------------------ */
void orc_glass_tile(const orc_params *p, const float *glass, int nglass, double glass_box,
                    float *pos, int *ids) {
  const int T = p->GlassTileFac;
  long long count = 0;
  for (int i = 0; i < T; i++)
    for (int j = 0; j < T; j++)
      for (int k = 0; k < T; k++)
        for (int n = 0; n < nglass; n++) {
          float x = glass[3 * n] / glass_box * (p->Box / T) + i * (p->Box / T);
          float y = glass[3 * n + 1] / glass_box * (p->Box / T) + j * (p->Box / T);
          float z = glass[3 * n + 2] / glass_box * (p->Box / T) + k * (p->Box / T);
          pos[3 * count] = x;
          pos[3 * count + 1] = y;
          pos[3 * count + 2] = z;
          if (ids) ids[count] = (int)(count + 1);
          count++;
        }
}

/* ------------------------------------------------------------------------ */
/* CIC readout and update, main.c:1440-1516, periodic_wrap 1534-1544         */
/* ------------------------------------------------------------------------ */
void orc_particles(const orc_params *p, const orc_cosmo *c, long long npart,
                   double *const disp[3], double *const disp2[3], float *pos, float *vel) {
  const int N = p->Nmesh, W = 2 * (N / 2 + 1);
  const double Box = p->Box;
  int nmesh3 = N * N * N;
#pragma omp parallel for schedule(static)
  for (long long n = 0; n < npart; n++) {
    float *P = pos + 3 * n, *V = vel + 3 * n;
    double u = P[0] / Box * N, v = P[1] / Box * N, w = P[2] / Box * N;
    int i = (int)u, j = (int)v, k = (int)w;
    if (i == N) i = N - 1;
    if (i < 0) i = 0;
    if (j == N) j = N - 1;
    if (k == N) k = N - 1;
    u -= i;
    v -= j;
    w -= k;
    int ii = i + 1, jj = j + 1, kk = k + 1;
    if (ii >= N) ii -= N; /* halo plane x=N holds plane 0 (main.c:1386-1430) */
    if (jj >= N) jj -= N;
    if (kk >= N) kk -= N;
    double f1 = (1 - u) * (1 - v) * (1 - w), f2 = (1 - u) * (1 - v) * (w);
    double f3 = (1 - u) * (v) * (1 - w), f4 = (1 - u) * (v) * (w);
    double f5 = (u) * (1 - v) * (1 - w), f6 = (u) * (1 - v) * (w);
    double f7 = (u) * (v) * (1 - w), f8 = (u) * (v) * (w);
#define AT(F, a, b, cc) F[((size_t)(a) * N + (b)) * W + (cc)]
    for (int ax = 0; ax < 3; ax++) {
      const double *D = disp[ax], *D2 = disp2[ax];
      double dis = AT(D, i, j, k) * f1 + AT(D, i, j, kk) * f2 + AT(D, i, jj, k) * f3 +
                   AT(D, i, jj, kk) * f4 + AT(D, ii, j, k) * f5 + AT(D, ii, j, kk) * f6 +
                   AT(D, ii, jj, k) * f7 + AT(D, ii, jj, kk) * f8;
      double dis2 = AT(D2, i, j, k) * f1 + AT(D2, i, j, kk) * f2 + AT(D2, i, jj, k) * f3 +
                    AT(D2, i, jj, kk) * f4 + AT(D2, ii, j, k) * f5 + AT(D2, ii, j, kk) * f6 +
                    AT(D2, ii, jj, k) * f7 + AT(D2, ii, jj, kk) * f8;
      dis2 /= (float)nmesh3;
      P[ax] += dis - 3. / 7. * dis2;
      V[ax] = dis * c->vel_prefac - 3. / 7. * dis2 * c->vel_prefac2;
      P[ax] = fmod(fmod(P[ax], Box) + Box, Box);
    }
#undef AT
  }
}

/* ------------------------------------------------------------------------ */
/* WDM thermal velocities of the snapshot writer, save.c:350-356 with         */
/* power.c:624-707: drand48 (never seeded, so from POSIX's default state),    */
/* the Fermi-Dirac table by Romberg, the reference's linear table search.     */
/* ------------------------------------------------------------------------ */
static uint64_t orc_d48 = 0x1234ABCD330EULL;
void orc_drand48_reset(void) { orc_d48 = 0x1234ABCD330EULL; }
double orc_drand48(void) {
  orc_d48 = (0x5DEECE66DULL * orc_d48 + 0xBULL) & ((1ULL << 48) - 1);
  return ldexp((double)orc_d48, -48);
}

#define ORC_FD_LEN 2000
static double fermi_dirac_kernel(const orc_ctx *cx, double x) { /* power.c:629-632 */
  (void)cx;
  return x * x / (exp(x) + 1);
}

void orc_wdm_thermal(const orc_params *p, long long npart, float *vel) {
  static double fvel[ORC_FD_LEN], fcum[ORC_FD_LEN];
  for (int i = 0; i < ORC_FD_LEN; i++) { /* power.c:638-646 */
    fvel[i] = 20.0 * i / (ORC_FD_LEN - 1.0);
    fcum[i] = nr_qromb(fermi_dirac_kernel, NULL, 0, fvel[i]);
  }
  for (int i = 0; i < ORC_FD_LEN; i++) fcum[i] /= fcum[ORC_FD_LEN - 1];
  double v0 = 0.012 * (1 + p->Redshift) * pow((p->Omega - p->OmegaBaryon) / 0.3, 1.0 / 3) *
              pow(p->HubbleParam / 0.65, 2.0 / 3) * pow(1.0 / p->WDM_PartMass_in_kev, 4.0 / 3);
  v0 *= 1.0e5 / p->UnitVelocity_in_cm_per_s;
  v0 *= sqrt(1 + p->Redshift);
  orc_drand48_reset();
  for (long long n = 0; n < npart; n++) {
    double pr = orc_drand48(); /* get_fermi_dirac_vel, power.c:667-683 */
    int i = 0;
    while (i < ORC_FD_LEN - 2)
      if (pr > fcum[i + 1])
        i++;
      else
        break;
    double u = (pr - fcum[i]) / (fcum[i + 1] - fcum[i]);
    double v = v0 * (fvel[i] * (1 - u) + fvel[i + 1] * u);
    double phi = 2 * M_PI * orc_drand48(); /* add_WDM_thermal_speeds, power.c:687-707 */
    double theta = acos(2 * orc_drand48() - 1);
    double vx = v * sin(theta) * cos(phi), vy = v * sin(theta) * sin(phi), vz = v * cos(theta);
    vel[3 * n] += vx;
    vel[3 * n + 1] += vy;
    vel[3 * n + 2] += vz;
  }
}

/* ------------------------------------------------------------------------ */
int orc_run(const orc_params *p, const double *alpha, const double *x0, const float *glass,
            int nglass, double glass_box, float *pos, float *vel, int *ids, double *phig_spec,
            double *phi_spec, double *phig_real, double *phing_real) {
  const int N = p->Nmesh;
  const size_t nd = 2 * (size_t)N * N * (N / 2 + 1);
  orc_cosmo c;
  int rc = orc_cosmology(p, &c);
  if (rc) return rc;
  uint32_t *st = malloc(sizeof(uint32_t) * N * N);
  double *cpot = malloc(sizeof(double) * nd);
  orc_seedtable(N, p->Seed, st);
  orc_gaussian(p, &c, st, cpot);
  free(st);
  if (phig_spec) memcpy(phig_spec, cpot, sizeof(double) * nd);
  rc = orc_ng_potential(p, &c, alpha, x0, cpot, phig_real, phing_real);
  if (rc) {
    free(cpot);
    return rc;
  }
  if (phi_spec) memcpy(phi_spec, cpot, sizeof(double) * nd);
  double *f[6];
  for (int a = 0; a < 6; a++) f[a] = malloc(sizeof(double) * nd);
  orc_lpt_fields(p, &c, cpot, f[0], f[1], f[2], f[3], f[4], f[5]);
  free(cpot);
  long long npart = (long long)p->GlassTileFac * p->GlassTileFac * p->GlassTileFac * nglass;
  orc_glass_tile(p, glass, nglass, glass_box, pos, ids);
  double *d1[3] = {f[0], f[1], f[2]}, *d2[3] = {f[3], f[4], f[5]};
  orc_particles(p, &c, npart, d1, d2, pos, vel);
  for (int a = 0; a < 6; a++) free(f[a]);
  return 0;
}


/* ------------------------------------------------------------ modal decomposition
 * Restatement of compute_mode_bispectra_covariance and compute_QS
 * (Aarambam/basis/tetrapyd.c:111-208, 210-281): the symmetrised mode product
 * with the same six-term order and the 1/6, weights applied to the first
 * factor, sums over points in increasing t. */
static double orc_sym_q(const double *ev, int nk, int p1, int p2, int p3, int i1, int i2, int i3) {
  const double m11 = ev[p1 * nk + i1], m12 = ev[p1 * nk + i2], m13 = ev[p1 * nk + i3];
  const double m21 = ev[p2 * nk + i1], m22 = ev[p2 * nk + i2], m23 = ev[p2 * nk + i3];
  const double m31 = ev[p3 * nk + i1], m32 = ev[p3 * nk + i2], m33 = ev[p3 * nk + i3];
  return (m11 * m22 * m33 + m11 * m23 * m32 + m12 * m21 * m33 + m12 * m23 * m31 + m13 * m21 * m32 +
          m13 * m22 * m31) / 6.0;
}

void orc_mode_covariance(double *cov, const double *w, const int *i1, const int *i2, const int *i3, int npts,
                         const int *p1, const int *p2, const int *p3, int n_modes, const double *evals, int pmax,
                         int nk) {
  (void)pmax;
#pragma omp parallel
  {
    double *qv = (double *)malloc(sizeof(double) * (size_t)npts);
#pragma omp for schedule(dynamic)
    for (int n = 0; n < n_modes; n++) {
      for (int t = 0; t < npts; t++) qv[t] = w[t] * orc_sym_q(evals, nk, p1[n], p2[n], p3[n], i1[t], i2[t], i3[t]);
      for (int m = 0; m < n_modes; m++) {
        double sum = 0.0;
        for (int t = 0; t < npts; t++) sum += orc_sym_q(evals, nk, p1[m], p2[m], p3[m], i1[t], i2[t], i3[t]) * qv[t];
        cov[(size_t)n * n_modes + m] = sum;
      }
    }
    free(qv);
  }
}

void orc_mode_qs(double *qs, const double *S, int n_shapes, const double *w, const int *i1, const int *i2,
                 const int *i3, int npts, const int *p1, const int *p2, const int *p3, int n_modes,
                 const double *evals, int pmax, int nk) {
  (void)pmax;
#pragma omp parallel
  {
    double *qv = (double *)malloc(sizeof(double) * (size_t)npts);
#pragma omp for schedule(dynamic)
    for (int n = 0; n < n_modes; n++) {
      for (int t = 0; t < npts; t++) qv[t] = w[t] * orc_sym_q(evals, nk, p1[n], p2[n], p3[n], i1[t], i2[t], i3[t]);
      for (int s = 0; s < n_shapes; s++) {
        double sum = 0.0;
        for (int t = 0; t < npts; t++) sum += S[(size_t)s * npts + t] * qv[t];
        qs[(size_t)s * n_modes + n] = sum;
      }
    }
    free(qv);
  }
}
