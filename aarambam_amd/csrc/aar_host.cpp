// Host side of the drop-in: parameter file, coefficient / transfer / glass
// readers, cosmology scalars, seed table and the Gadget, potential and
// input-spectrum writers.  Semantics follow Aarambam/LPT/Basis for one
// process holding the whole grid (NTask = NTaskWithN = 1).
#include <math.h>
#include <cmath>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <memory>
#include <thread>
#include <vector>

#include "../../include/aarambam.h"
#include "aar_common.h"
#include "aar_ranlxd.h"

static const double kGravity = 6.672e-8;    // allvars.h:5
static const double kHubble = 3.2407789e-18;  // allvars.h:6
static const double kPi = AAR_PI;

extern "C" const char* aar_status_string(int s) {
  switch (s) {
    case AAR_OK: return "ok";
    case AAR_ERR_ALLOC: return "allocation failed (FatalError 1)";
    case 2: return "incorrect glass header size (FatalError 2)";
    case 3: return "incorrect glass position block (FatalError 3)";
    case 10: return "cannot write output file (FatalError 10)";
    case AAR_ERR_TABLE_FILE: return "cannot read coefficient or transfer table (FatalError 17)";
    case 121: return "glass file not found (FatalError 121)";
    case 777: return "duplicate k in transfer table (FatalError 777)";
    case 778: return "read error (FatalError 778)";
    case AAR_ERR_KLOWER: return "transfer table klower too large (FatalError 1111)";
    case AAR_ERR_WRITE_PARALLEL:
      return "NumFilesWrittenInParallel exceeds the number of tasks (FatalError 24131)";
    case AAR_ERR_BAD_ARG: return "invalid argument";
    case AAR_ERR_HIP: return "HIP runtime error";
    case AAR_ERR_HBM_BUDGET: return "job does not fit the HBM budget";
    case AAR_ERR_NMAX: return "Nmax < Nsample (main.c:200 assert)";
    case AAR_ERR_UNSUPPORTED: return "configuration not supported";
    case AAR_ERR_NO_DEVICE: return "no HIP device";
    case AAR_ERR_SCRATCH: return "job needs more device memory than the fixed 64 MiB scratch of aar_job_run_noalloc";
    case 100010: return "parameter file error";
    default: return "unknown status";
  }
}

// ------------------------------------------------------------------ parameter file
// read_parameterfile of Aarambam/LPT/Basis/read_param.c:8-262; with res the
// tag set of Aarambam/LPT/ResBasis/read_param.c:8-274 (A_res, w_res, phi_res
// follow SavePotentialField, which fixes the order of the missing-tag report).
static int read_params(const char* path, aar_params* out, char* msg, size_t msglen, bool res) {
  if (!path || !out) return AAR_ERR_BAD_ARG;
  enum { F, S, I };
  struct Tag {
    const char* name;
    int type;
    void* addr;
    bool seen;
  };
  aar_params p;
  memset(&p, 0, sizeof(p));
  std::vector<Tag> tags = {
      {"Omega", F, &p.Omega},
      {"OmegaLambda", F, &p.OmegaLambda},
      {"OmegaBaryon", F, &p.OmegaBaryon},
      {"OmegaDM_2ndSpecies", F, &p.OmegaDM_2ndSpecies},
      {"HubbleParam", F, &p.HubbleParam},
      {"ShapeGamma", F, &p.ShapeGamma},
      {"Sigma8", F, &p.Sigma8},
      {"PrimordialIndex", F, &p.PrimordialIndex},
      {"Box", F, &p.Box},
      {"Redshift", F, &p.Redshift},
      {"Fnl", F, &p.Fnl},
      {"N_modes", I, &p.N_modes},
      {"FileWithAlphaCoeff", S, p.FileWithAlphaCoeff},
      {"FileWithX0Coeff", S, p.FileWithX0Coeff},
      {"SavePotentialField", I, &p.SavePotentialField},
      {"A_res", F, &p.A_res},
      {"w_res", F, &p.w_res},
      {"phi_res", F, &p.phi_res},
      {"FixedAmplitude", I, &p.FixedAmplitude},
      {"PhaseFlip", I, &p.PhaseFlip},
      {"Nmesh", I, &p.Nmesh},
      {"Nsample", I, &p.Nsample},
      {"Nmax", I, &p.Nmax},
      {"GlassFile", S, p.GlassFile},
      {"FileWithInputSpectrum", S, p.FileWithInputSpectrum},
      {"FileWithInputTransfer", S, p.FileWithInputTransfer},
      {"GlassTileFac", I, &p.GlassTileFac},
      {"Seed", I, &p.Seed},
      {"SphereMode", I, &p.SphereMode},
      {"NumFilesWrittenInParallel", I, &p.NumFilesWrittenInParallel},
      {"OutputDir", S, p.OutputDir},
      {"FileBase", S, p.FileBase},
      {"WhichSpectrum", I, &p.WhichSpectrum},
      {"WhichTransfer", I, &p.WhichTransfer},
      {"UnitVelocity_in_cm_per_s", F, &p.UnitVelocity_in_cm_per_s},
      {"UnitLength_in_cm", F, &p.UnitLength_in_cm},
      {"UnitMass_in_g", F, &p.UnitMass_in_g},
      {"InputSpectrum_UnitLength_in_cm", F, &p.InputSpectrum_UnitLength_in_cm},
      {"WDM_On", I, &p.WDM_On},
      {"WDM_Vtherm_On", I, &p.WDM_Vtherm_On},
      {"WDM_PartMass_in_kev", F, &p.WDM_PartMass_in_kev},
  };
  if (!res)  // Basis: the resonance tags are "not allowed"
    tags.erase(std::remove_if(tags.begin(), tags.end(), [&](const Tag& t) { return t.addr == &p.A_res ||
                                t.addr == &p.w_res || t.addr == &p.phi_res; }),
               tags.end());
  std::string report;
  bool error = false;
  FILE* fd = fopen(path, "r");
  if (!fd) {
    report += std::string("Parameter file ") + path + " not found.\n";
    error = true;
  } else {
    // read_param.c:186-229: lines of at most 199 characters, first two
    // whitespace-separated tokens, '%' comments; each tag exactly once.
    char buf[200], b1[200], b2[200], b3[200];
    while (!feof(fd)) {
      buf[0] = 0;
      if (!fgets(buf, 200, fd)) buf[0] = 0;
      if (sscanf(buf, "%199s%199s%199s", b1, b2, b3) < 2) continue;
      if (b1[0] == '%') continue;
      Tag* hit = nullptr;
      for (auto& t : tags)
        if (!t.seen && strcmp(b1, t.name) == 0) {
          hit = &t;
          break;
        }
      if (!hit) {
        report += std::string("Error in file ") + path + ":   Tag '" + b1 +
                  "' not allowed or multiple defined.\n";
        error = true;
        continue;
      }
      hit->seen = true;
      if (hit->type == F) *(double*)hit->addr = atof(b2);
      else if (hit->type == I) *(int*)hit->addr = atoi(b2);
      else {
        strncpy((char*)hit->addr, b2, AAR_STR - 1);
        ((char*)hit->addr)[AAR_STR - 1] = 0;
      }
    }
    fclose(fd);
  }
  for (auto& t : tags)
    if (!t.seen) {
      report += std::string("Error. I miss a value for tag '") + t.name + "' in parameter file '" +
                path + "'.\n";
      error = true;
    }
  if (msg && msglen) {
    strncpy(msg, report.c_str(), msglen - 1);
    msg[msglen - 1] = 0;
  }
  p.Variant = res ? 1 : 0;
  *out = p;
  return error ? 100010 : AAR_OK;
}

extern "C" int aar_read_params(const char* path, aar_params* out, char* msg, size_t msglen) {
  return read_params(path, out, msg, msglen, false);
}

extern "C" int aar_read_params_res(const char* path, aar_params* out, char* msg, size_t msglen) {
  return read_params(path, out, msg, msglen, true);
}

// ------------------------------------------------------------------ tables
extern "C" int aar_load_alpha(const char* path, int nm, double* a) {
  if (!path || !a || nm < 3) return AAR_ERR_BAD_ARG;
  FILE* fd = fopen(path, "r");
  if (!fd) return AAR_ERR_TABLE_FILE;
  std::fill(a, a + (size_t)nm * nm * nm, 0.0);
  int n1, n2, n3;
  double v;
  int rc = AAR_OK;
  auto at = [&](int x, int y, int z) -> double& { return a[((size_t)x * nm + y) * nm + z]; };
  while (fscanf(fd, " %d %d %d %lf ", &n1, &n2, &n3, &v) == 4) {  // main.c:231
    if (n1 < 0 || n2 < 0 || n3 < 0 || n1 >= nm || n2 >= nm || n3 >= nm) {
      rc = AAR_ERR_BAD_ARG;  // the reference writes outside its table here
      break;
    }
    at(n3, n2, n1) = v;
    at(n3, n1, n2) = v;
    at(n2, n3, n1) = v;
    at(n2, n1, n3) = v;
    at(n1, n3, n2) = v;
    at(n1, n2, n3) = v;
  }
  fclose(fd);
  return rc;
}

extern "C" int aar_load_x0(const char* path, int nm, double* x0) {
  if (!path || !x0 || nm < 1) return AAR_ERR_BAD_ARG;
  FILE* fd = fopen(path, "r");
  if (!fd) return AAR_ERR_TABLE_FILE;
  std::fill(x0, x0 + 2 * (size_t)nm, 0.0);
  int n;
  double a, b;
  int rc = AAR_OK;
  while (fscanf(fd, " %d %lf %lf ", &n, &a, &b) == 3) {  // main.c:274-285
    if (n < 0 || n >= nm) continue;  // rows beyond N_modes are never read by the sum
    x0[2 * n] = a;
    x0[2 * n + 1] = b;
  }
  fclose(fd);
  return rc;
}

extern "C" int aar_load_transfer(const char* path, int* nrows, double* k, double* t) {
  if (!path || !nrows) return AAR_ERR_BAD_ARG;
  FILE* fd = fopen(path, "r");
  if (!fd) return AAR_ERR_TABLE_FILE;
  double a, b;
  int n = 0;
  while (fscanf(fd, " %le %le ", &a, &b) == 2) {  // power.c:85-109
    if (k && t) {
      if (n >= *nrows) break;
      k[n] = a;
      t[n] = b;
    }
    ++n;
  }
  fclose(fd);
  *nrows = n;
  return AAR_OK;
}

// Gadget-1 glass, possibly split over name.0, name.1, ... (read_glass.c:25-88, 226-260).
extern "C" int aar_read_glass(const char* path, int* nglass, float* pos, double* box,
                              uint8_t* header_out) {
  if (!path || !nglass) return AAR_ERR_BAD_ARG;
  std::string base(path);
  int numfiles = 0;
  uint8_t hdr[256];
  {
    std::string f0 = base + ".0";
    FILE* fd = fopen(f0.c_str(), "rb");
    int dummy;
    if (fd) {
      if (fread(&dummy, 4, 1, fd) != 1 || fread(hdr, 256, 1, fd) != 1) {
        fclose(fd);
        return 778;
      }
      fclose(fd);
      memcpy(&numfiles, hdr + 124, 4);  // num_files
    } else if ((fd = fopen(base.c_str(), "rb"))) {
      fclose(fd);
      numfiles = 1;
    } else {
      return 121;
    }
  }
  long long total = 0, filled = 0;
  for (int num = 0; num < numfiles; ++num) {
    std::string fn = numfiles > 1 ? base + "." + std::to_string(num) : base;
    FILE* fd = fopen(fn.c_str(), "rb");
    if (!fd) return 1;
    unsigned d1 = 0, d2 = 0;
    uint8_t h[256];
    if (fread(&d1, 4, 1, fd) != 1 || fread(h, 256, 1, fd) != 1 || fread(&d2, 4, 1, fd) != 1) {
      fclose(fd);
      return 778;
    }
    if (d1 != 256 || d2 != 256) {
      fclose(fd);
      return 2;
    }
    unsigned npart[6], nall[6];
    memcpy(npart, h, 24);
    memcpy(nall, h + 96, 24);
    long long nlocal = 0;
    for (int q = 0; q < 6; ++q) nlocal += npart[q];
    if (num == 0) {
      total = 0;
      for (int q = 0; q < 6; ++q) total += nall[q];
      if (box) memcpy(box, h + 128, 8);
      if (header_out) memcpy(header_out, h, 256);
      if (!pos) {
        fclose(fd);
        *nglass = (int)total;
        return AAR_OK;
      }
      if (total > *nglass) {
        fclose(fd);
        return AAR_ERR_BAD_ARG;
      }
    }
    if (fread(&d1, 4, 1, fd) != 1) {
      fclose(fd);
      return 778;
    }
    if (filled + nlocal > total ||
        fread(pos + 3 * filled, sizeof(float), 3 * nlocal, fd) != (size_t)(3 * nlocal) ||
        fread(&d2, 4, 1, fd) != 1) {
      fclose(fd);
      return 778;
    }
    if (d1 != 12 * nlocal || d2 != 12 * nlocal) {
      fclose(fd);
      return 3;
    }
    filled += nlocal;
    fclose(fd);
  }
  *nglass = (int)total;
  return AAR_OK;
}

// ------------------------------------------------------------------ cosmology
namespace {
struct Cosmo {
  const aar_params& p;
  const aar_tables* t;
  double norm = 1.0, r_tophat = 0.0;
  std::vector<double> logk, logT;
  int status = AAR_OK;

  Cosmo(const aar_params& pp, const aar_tables* tt) : p(pp), t(tt) {
    if (p.WhichTransfer == 2) {  // read_transfer_table, power.c:111-131
      if (!t || t->ntransfer < 2 || !t->transfer_k || !t->transfer_t) {
        status = AAR_ERR_TABLE_FILE;
        return;
      }
      std::vector<std::pair<double, double>> rows;
      for (int i = 0; i < t->ntransfer; ++i)
        rows.push_back({log10(t->transfer_k[i]), log10(t->transfer_t[i])});
      std::stable_sort(rows.begin(), rows.end(),
                       [](const std::pair<double, double>& a, const std::pair<double, double>& b) {
                         return a.first < b.first;
                       });
      if (rows[0].first >= -4.6) {
        status = AAR_ERR_KLOWER;
        return;
      }
      for (size_t i = 1; i < rows.size(); ++i)
        if (rows[i].first == rows[i - 1].first) status = 777;
      for (auto& r : rows) {
        logk.push_back(r.first);
        logT.push_back(r.second - rows[0].second);
      }
    }
  }

  double tk_eh(double k) const {  // power.c:519-548
    double hubble = p.HubbleParam, omegam = p.Omega;
    double ombh2 = p.OmegaBaryon * p.HubbleParam * p.HubbleParam;
    if (p.OmegaBaryon == 0) ombh2 = 0.04 * p.HubbleParam * p.HubbleParam;
    k *= (3.085678e24 / p.UnitLength_in_cm);
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

  double transfer(double k) const {  // power.c:32-63
    if (p.WhichTransfer == 1) return tk_eh(k);
    if (p.WhichTransfer == 2) {  // power.c:181-217
      k *= (p.InputSpectrum_UnitLength_in_cm / p.UnitLength_in_cm);
      double lk = log10(k);
      if (lk < logk.front() || lk > logk.back()) return 0;
      int lo = 0, hi = (int)logk.size() - 1;
      while (hi - lo > 1) {
        int mid = (hi + lo) / 2;
        if (lk < logk[mid]) hi = mid;
        else lo = mid;
      }
      double u = (lk - logk[lo]) / (logk[hi] - logk[lo]);
      return pow(10.0, (1 - u) * logT[lo] + u * logT[hi]);
    }
    return 1.0;  // tk_Efstathiou
  }

  double powerspec(double k) const {  // power.c:240-288
    double power;
    if (p.Variant == 1) {  // ResBasis/power.c:240-260: oscillating primordial P(k) times T(k)^2
      power = norm * exp(p.PrimordialIndex * log(k));
      if (p.WhichSpectrum == 1) power *= 1 + p.A_res * cos(p.w_res * k + p.phi_res);
      if (p.WhichSpectrum == 2) power *= 1 + p.A_res * cos(p.w_res * log(k) + p.phi_res);
      power *= transfer(k) * transfer(k);
    } else if (p.WhichSpectrum == 0) {
      power = norm * exp(p.PrimordialIndex * log(k)) * transfer(k) * transfer(k);
    } else {  // 1: Eisenstein & Hu spectrum
      power = norm * k * pow(tk_eh(k), 2);
      power *= pow(k, p.PrimordialIndex - 1.0);
    }
    if (p.WDM_On == 1) {
      double alpha = 0.048 * pow((p.Omega - p.OmegaBaryon) / 0.4, 0.15) *
                     pow(p.HubbleParam / 0.65, 1.3) * pow(1.0 / p.WDM_PartMass_in_kev, 1.15);
      double Tf = pow(1 + pow(alpha * k * (3.085678e24 / p.UnitLength_in_cm), 2 * 1.2), -5.0 / 1.2);
      power *= Tf * Tf;
    }
    return power;
  }

  double sigma2_int(double k) const {  // power.c:561-576
    double kr = r_tophat * k, kr2 = kr * kr, kr3 = kr2 * kr;
    if (kr < 1e-8) return 0;
    double w = 3 * (sin(kr) / kr3 - cos(kr) / kr2);
    return 4 * kPi * k * k * w * w * powerspec(k);
  }

  double growth_int(double a) const {  // power.c:595-598
    return pow(a / (p.Omega + (1 - p.Omega - p.OmegaLambda) * a + p.OmegaLambda * a * a * a), 1.5);
  }

  // Romberg integration of Numerical Recipes (nrsrc/qromb.c, trapzd.c,
  // polint.c): trapezoid refinements, K=5 point polynomial extrapolation to
  // h=0, stop when |error estimate| <= 1e-8 |value|, at most 40 steps.
  template <class Fn> static double romberg(Fn f, double a, double b) {
    double s[42], h[43];
    double st = 0;
    h[1] = 1.0;
    for (int j = 1; j <= 40; ++j) {
      if (j == 1) {
        st = 0.5 * (b - a) * (f(a) + f(b));
      } else {
        int it = 1;
        for (int q = 1; q < j - 1; ++q) it <<= 1;
        double tnm = it, del = (b - a) / tnm, x = a + 0.5 * del, sum = 0.0;
        for (int q = 1; q <= it; ++q, x += del) sum += f(x);
        st = 0.5 * (st + (b - a) * sum / tnm);
      }
      s[j] = st;
      if (j >= 5) {
        double y, dy;
        neville(&h[j - 5], &s[j - 5], 5, 0.0, &y, &dy);
        if (fabs(dy) <= 1.0e-8 * fabs(y)) return y;
      }
      h[j + 1] = 0.25 * h[j];
    }
    fprintf(stderr, "Too many steps in routine qromb\n");
    return 0.0;
  }

  static void neville(const double xa[], const double ya[], int n, double x, double* y, double* dy) {
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
        double den = (c[i + 1] - d[i]) / (ho - hp);
        d[i] = hp * den;
        c[i] = ho * den;
      }
      *y += (*dy = (2 * ns < (n - m) ? c[ns + 1] : d[ns--]));
    }
  }

  double growth(double a) const {  // power.c:585-592
    double ha = sqrt(p.Omega / (a * a * a) + (1 - p.Omega - p.OmegaLambda) / (a * a) + p.OmegaLambda);
    return ha * romberg([this](double x) { return growth_int(x); }, 0, a);
  }
  double growth_factor(double a0, double a1) const { return growth(a1) / growth(a0); }

  double f_omega(double a) const {
    double om = p.Omega / (p.Omega + a * (1 - p.Omega - p.OmegaLambda) + a * a * a * p.OmegaLambda);
    return pow(om, 0.6);
  }
  double f2_omega(double a) const {
    double om = p.Omega / (p.Omega + a * (1 - p.Omega - p.OmegaLambda) + a * a * a * p.OmegaLambda);
    return 2 * pow(om, 4. / 7.);
  }
};
}  // namespace

extern "C" int aar_cosmology(const aar_params* p, const aar_tables* t, aar_cosmo* c) {
  if (!p || !c) return AAR_ERR_BAD_ARG;
  // Basis: 0 (T(k) from WhichTransfer) and 1 (Eisenstein & Hu); ResBasis: 0, 1, 2
  // (its PowerSpec leaves every other value undefined, ResBasis/power.c:244-258).
  const int ws_max = p->Variant == 1 ? 2 : 1;
  if (p->Variant != 0 && p->Variant != 1) return AAR_ERR_BAD_ARG;
  if (p->WhichSpectrum < 0 || p->WhichSpectrum > ws_max) return AAR_ERR_UNSUPPORTED;
  if (p->WhichTransfer == 3) return 117;  // power.c:45-50
  Cosmo cm(*p, t);
  if (cm.status) return cm.status;
  memset(c, 0, sizeof(*c));
  c->UnitTime_in_s = p->UnitLength_in_cm / p->UnitVelocity_in_cm_per_s;
  c->G = kGravity / pow(p->UnitLength_in_cm, 3) * p->UnitMass_in_g * pow(c->UnitTime_in_s, 2);
  c->Hubble = kHubble * c->UnitTime_in_s;
  c->InitTime = 1 / (1 + p->Redshift);
  const double R8 = 8 * (3.085678e24 / p->UnitLength_in_cm);
  cm.norm = 1.0;
  cm.r_tophat = R8;
  const double res = cm.romberg([&cm](double k) { return cm.sigma2_int(k); }, 0, 500.0 * 1 / R8);
  c->sigma8_in = sqrt(res);
  c->Norm = p->Sigma8 * p->Sigma8 / res;
  c->D0 = cm.growth_factor(1.e-8, 1.0) * 1.e-8;
  c->Dplus = cm.growth_factor(c->InitTime, 1.0);
  // Basis sets Anorm for WhichSpectrum 0 only (power.c:426-446); ResBasis always (power.c:418)
  c->Anorm = (p->Variant == 1 || p->WhichSpectrum == 0) ? c->Norm : 0.0;
  const double a = c->InitTime;
  c->hubble_a = c->Hubble * sqrt(p->Omega / pow(a, 3) + (1 - p->Omega - p->OmegaLambda) / pow(a, 2) +
                                 p->OmegaLambda);
  c->vel_prefac = a * c->hubble_a * cm.f_omega(a) / sqrt(a);
  c->vel_prefac2 = a * c->hubble_a * cm.f2_omega(a) / sqrt(a);
  c->fac = pow(2 * kPi / p->Box, 1.5);
  c->Beta = 1.5 * p->Omega /
            (2998. * 2998. / p->UnitLength_in_cm / p->UnitLength_in_cm * 3.085678e24 * 3.085678e24) /
            c->D0;
  return AAR_OK;
}

// Host evaluation of T(k) for the device tables (used by aar_job.cpp).
double aar_host_transfer(const aar_params* p, const aar_tables* t, double k) {
  Cosmo cm(*p, t);
  return cm.transfer(k);
}

int aar_host_transfer_tables(const aar_params* p, const aar_tables* t, std::vector<double>& logk,
                             std::vector<double>& logT) {
  Cosmo cm(*p, t);
  if (cm.status) return cm.status;
  logk = cm.logk;
  logT = cm.logT;
  return AAR_OK;
}

// ------------------------------------------------------------------ ranlxd jump-ahead
// The subtract-with-borrow recursion x_n = x_{n-5} - x_{n-12} - c_{n-1} (mod b = 2^48) is a
// multiplicative congruential generator in another form (Marsaglia and Zaman 1991; Tezuka,
// L'Ecuyer and Couture 1993).  With m = b^12 - b^5 + 1 and
//   X_n = sum_{j<12} x_{n-12+j} b^j - sum_{j<5} x_{n-5+j} b^j + c_{n-1},
// the recursion gives b X_{n+1} = X_n + x_n m exactly, so X_{n+k} = b^-k X_n (mod m), and the
// outputs follow from the true X: x_n = -X_n mod b (m = 1 mod b).  X_n lies in (-b^5, b^12],
// so its residue fixes it except within b^5 of a multiple of m, where the jump declines.
namespace {
constexpr int kBigLimbs = 20;  // 1280 bits: products of two residues below 2^640
struct Big {
  uint64_t w[kBigLimbs] = {};
};
int big_cmp(const Big& a, const Big& b) {
  for (int i = kBigLimbs - 1; i >= 0; --i)
    if (a.w[i] != b.w[i]) return a.w[i] < b.w[i] ? -1 : 1;
  return 0;
}
void big_add(Big& a, const Big& b) {
  unsigned __int128 c = 0;
  for (int i = 0; i < kBigLimbs; ++i) {
    c += (unsigned __int128)a.w[i] + b.w[i];
    a.w[i] = (uint64_t)c;
    c >>= 64;
  }
}
void big_sub(Big& a, const Big& b) {  // a >= b
  uint64_t br = 0;
  for (int i = 0; i < kBigLimbs; ++i) {
    const uint64_t x = a.w[i], y = b.w[i];
    const uint64_t d = x - y - br;
    br = (x < y) || (x - y < br);
    a.w[i] = d;
  }
}
Big big_shl(const Big& a, int bits) {
  Big r;
  const int q = bits / 64, s = bits % 64;
  for (int i = kBigLimbs - 1; i >= q; --i) {
    uint64_t v = a.w[i - q] << s;
    if (s && i - q - 1 >= 0) v |= a.w[i - q - 1] >> (64 - s);
    r.w[i] = v;
  }
  return r;
}
Big big_shr(const Big& a, int bits) {
  Big r;
  const int q = bits / 64, s = bits % 64;
  for (int i = 0; i + q < kBigLimbs; ++i) {
    uint64_t v = a.w[i + q] >> s;
    if (s && i + q + 1 < kBigLimbs) v |= a.w[i + q + 1] << (64 - s);
    r.w[i] = v;
  }
  return r;
}
Big big_low(const Big& a, int bits) {  // a mod 2^bits
  Big r = a;
  const int q = bits / 64, s = bits % 64;
  for (int i = q; i < kBigLimbs; ++i) r.w[i] = (i == q && s) ? (r.w[i] & ((1ull << s) - 1)) : 0;
  return r;
}
Big big_mul(const Big& a, const Big& b) {  // low 1280 bits of a b
  Big r;
  for (int i = 0; i < kBigLimbs; ++i) {
    if (!a.w[i]) continue;
    unsigned __int128 c = 0;
    for (int j = 0; i + j < kBigLimbs; ++j) {
      c += (unsigned __int128)a.w[i] * b.w[j] + r.w[i + j];
      r.w[i + j] = (uint64_t)c;
      c >>= 64;
    }
  }
  return r;
}
Big big_digits(const uint64_t* d, int n) {  // sum_j d[j] 2^(48 j), d[j] < 2^48
  Big r;
  for (int j = n - 1; j >= 0; --j) {
    r = big_shl(r, 48);
    r.w[0] |= d[j];
  }
  return r;
}
const Big& swb_m() {  // 2^576 - 2^240 + 1
  static const Big m = [] {
    Big one, t;
    one.w[0] = 1;
    t = big_shl(one, 576);
    big_sub(t, big_shl(one, 240));
    big_add(t, one);
    return t;
  }();
  return m;
}
void big_reduce(Big& p) {  // p mod m, folding 2^576 = 2^240 - 1 (mod m)
  for (Big h = big_shr(p, 576); big_cmp(h, Big()) != 0; h = big_shr(p, 576)) {
    Big t = big_low(p, 576);
    big_add(t, big_shl(h, 240));
    big_sub(t, h);
    p = t;
  }
  while (big_cmp(p, swb_m()) >= 0) big_sub(p, swb_m());
}
Big big_mulmod(const Big& a, const Big& b) {
  Big p = big_mul(a, b);
  big_reduce(p);
  return p;
}
// b^-k mod m, b = 2^48: b^-1 = m - (m - 1) / b
Big swb_inv_pow(uint64_t k) {
  Big mm1 = swb_m(), one;
  one.w[0] = 1;
  big_sub(mm1, one);
  Big base = swb_m();
  big_sub(base, big_shr(mm1, 48));
  Big r = one;
  while (k) {
    if (k & 1) r = big_mulmod(r, base);
    base = big_mulmod(base, base);
    k >>= 1;
  }
  return r;
}
}  // namespace

// The generator `steps` (>= 12) recursion steps further on, registers oldest first as after
// ranlxd_update; false (r unchanged) where the jump declines.
bool aar_ranlxd_jump(Ranlxd& r, uint64_t steps) {
  if (steps < 12) return false;
  const Big& m = swb_m();
  // residue of X_n = A - B + c
  Big x = big_digits(r.s, 12);
  Big c;
  c.w[0] = r.carry;
  big_add(x, c);
  big_reduce(x);
  const Big bb = big_digits(r.s + 7, 5);
  if (big_cmp(x, bb) < 0) big_add(x, m);
  big_sub(x, bb);
  // residue of X_{n+steps-12}, then the true value
  x = big_mulmod(x, swb_inv_pow(steps - 12));
  Big lo, hi = m;
  lo.w[0] = 1;
  lo = big_shl(lo, 240);
  big_sub(hi, lo);
  if (big_cmp(x, lo) < 0 || big_cmp(x, hi) > 0) return false;
  // twelve outputs x_n = -X_n mod b, X_{n+1} = (X_n + x_n m) / b
  uint64_t s[12];
  for (int j = 0; j < 12; ++j) {
    const uint64_t xn = (0 - x.w[0]) & kMask48;
    s[j] = xn;
    Big t = m;
    Big k;
    k.w[0] = xn;
    t = big_mul(t, k);
    big_add(x, t);
    x = big_shr(x, 48);
  }
  // the borrow: c = X - A + B, 0 or 1
  big_add(x, big_digits(s + 7, 5));
  const Big a = big_digits(s, 12);
  if (big_cmp(x, a) < 0) return false;
  big_sub(x, a);
  Big one;
  one.w[0] = 1;
  if (big_cmp(x, one) > 0) return false;
  for (int j = 0; j < 12; ++j) r.s[j] = s[j];
  r.carry = x.w[0];
  return true;
}

// ------------------------------------------------------------------ seed table
// Draws of shells [i0, i1) of the fill order of main.c:321-349; shell i takes 8 i + 4 draws, so
// shell i starts at draw 4 i^2.
static void seed_shells(int N, uint32_t seed, int i0, int i1, uint32_t* st) {
  RanlxdStream g;
  g.seed(seed);
  const uint64_t d0 = 4ull * i0 * i0;
  if (d0) {  // the state after d0 / 12 + 1 updates, at output d0 % 12
    const uint64_t u = d0 / 12 + 1;
    if (!aar_ranlxd_jump(g.r, 202 * u))
      for (uint64_t q = 0; q < u; ++q) ranlxd_update(g.r);
    g.pos = (int)(d0 % 12);
  }
  auto draw = [&]() { return (uint32_t)(0x7fffffff * g.uniform()); };
  for (int i = i0; i < i1; i++) {
    for (int j = 0; j < i; j++) st[i * N + j] = draw();
    for (int j = 0; j < i + 1; j++) st[j * N + i] = draw();
    for (int j = 0; j < i; j++) st[(N - 1 - i) * N + j] = draw();
    for (int j = 0; j < i + 1; j++) st[(N - 1 - j) * N + i] = draw();
    for (int j = 0; j < i; j++) st[i * N + (N - 1 - j)] = draw();
    for (int j = 0; j < i + 1; j++) st[j * N + (N - 1 - i)] = draw();
    for (int j = 0; j < i; j++) st[(N - 1 - i) * N + (N - 1 - j)] = draw();
    for (int j = 0; j < i + 1; j++) st[(N - 1 - j) * N + (N - 1 - i)] = draw();
  }
}

// One serial stream of N^2 draws, filled by up to eight threads that each jump the generator to
// the first draw of their shells (equal numbers of draws: shell boundaries at h sqrt(t / T)).
extern "C" int aar_seedtable_threads(int N, int seed, uint32_t* st, int threads) {
  if (N < 2 || !st || threads < 1) return AAR_ERR_BAD_ARG;
  const int h = N / 2;
  const int T = std::min(threads, std::max(1, h / 16));
  if (T == 1) {
    seed_shells(N, (uint32_t)seed, 0, h, st);
    return AAR_OK;
  }
  std::vector<std::thread> pool;
  int i0 = 0;
  for (int t = 1; t <= T; ++t) {
    const int i1 = t == T ? h : std::min(h, (int)std::lround(h * std::sqrt((double)t / T)));
    if (i1 > i0) pool.emplace_back(seed_shells, N, (uint32_t)seed, i0, i1, st);
    i0 = std::max(i0, i1);
  }
  for (auto& th : pool) th.join();
  return AAR_OK;
}

extern "C" int aar_seedtable(int N, int seed, uint32_t* st) {
  const unsigned hw = std::thread::hardware_concurrency();
  return aar_seedtable_threads(N, seed, st, N >= 512 ? (int)std::min(8u, std::max(1u, hw)) : 1);
}

extern "C" unsigned long aar_ranlxd_nth(unsigned long seed, int luxury, int n) {
  // GSL rng_test semantics: the n-th gsl_rng_get = (unsigned long)(u * 2^32),
  // with the generator's register bookkeeping of rng/ranlxd.c (ir, jr, ir_old).
  Ranlxd r;
  ranlxd_seed(r, (uint32_t)seed);
  int ir = 11, jr = 7, ir_old = 0;
  unsigned long k = 0;
  for (int i = 0; i < n; ++i) {
    ir = (ir + 1) % 12;
    if (ir == ir_old) {
      for (int q = 0; q < luxury; ++q) {
        ranlxd_step(r.s[ir], r.s[jr], r.carry);
        ir = (ir + 1) % 12;
        jr = (jr + 1) % 12;
      }
      ir_old = ir;
    }
    k = (unsigned long)((double)r.s[ir] * kTwoM48 * 4294967296.0);
  }
  return k;
}

// ------------------------------------------------------------------ WDM thermal velocities
// With WDM_On 1 and WDM_Vtherm_On 1 the reference's snapshot writer adds a random thermal
// velocity to every particle's velocity as it writes it (save.c:350-356; the particle array
// itself keeps the ZA + 2LPT velocities), drawn by power.c:624-707 from drand48, which the
// reference never seeds: every writing task starts from drand48's default state.
namespace {
// drand48 (POSIX): X <- (0x5DEECE66D X + 0xB) mod 2^48 from X = 0x1234ABCD330E; value X 2^-48.
struct Drand48 {
  uint64_t x = 0x1234ABCD330EULL;
  double next() {
    x = (0x5DEECE66DULL * x + 0xBULL) & ((1ULL << 48) - 1);
    return ldexp((double)x, -48);
  }
};

// fermi_dirac_init (power.c:634-663): the cumulative distribution of x^2 / (e^x + 1) on 2000
// speeds in [0, 20] by Romberg integration, and WDM_V0 in internal units.
struct FermiDirac {
  static constexpr int kLen = 2000;
  double vel[kLen], cum[kLen], cummax[kLen];
  double v0 = 0;
  explicit FermiDirac(const aar_params* p) {
    for (int i = 0; i < kLen; i++) {
      vel[i] = 20.0 * i / (kLen - 1.0);
      cum[i] = Cosmo::romberg([](double x) { return x * x / (exp(x) + 1); }, 0, vel[i]);
    }
    for (int i = 0; i < kLen; i++) cum[i] /= cum[kLen - 1];
    // the running maximum: power.c's linear search stops at the first i with p <= cum[i + 1],
    // which is also the first i with p <= max(cum[1..i + 1]), so a binary search on it agrees
    cummax[0] = cum[0];
    for (int i = 1; i < kLen; i++) cummax[i] = std::max(cummax[i - 1], cum[i]);
    v0 = 0.012 * (1 + p->Redshift) * pow((p->Omega - p->OmegaBaryon) / 0.3, 1.0 / 3) *
         pow(p->HubbleParam / 0.65, 2.0 / 3) * pow(1.0 / p->WDM_PartMass_in_kev, 4.0 / 3);
    printf("\nWarm dark matter rms velocity dispersion at starting redshift = %g km/sec\n\n", 3.59714 * v0);
    v0 *= 1.0e5 / p->UnitVelocity_in_cm_per_s;
    v0 *= sqrt(1 + p->Redshift);  // peculiar to Gadget's cosmological velocity
  }
  double speed(Drand48& r) const {  // get_fermi_dirac_vel, power.c:667-683
    const double pr = r.next();
    int i = (int)(std::lower_bound(cummax + 1, cummax + kLen - 1, pr) - (cummax + 1));
    if (i > kLen - 2) i = kLen - 2;
    const double u = (pr - cum[i]) / (cum[i + 1] - cum[i]);
    return vel[i] * (1 - u) + vel[i + 1] * u;
  }
  void kick(float* v3, Drand48& r) const {  // add_WDM_thermal_speeds, power.c:687-707
    const double v = v0 * speed(r);
    const double phi = 2 * M_PI * r.next();
    const double theta = acos(2 * r.next() - 1);
    v3[0] += v * sin(theta) * cos(phi);
    v3[1] += v * sin(theta) * sin(phi);
    v3[2] += v * cos(theta);
  }
};

bool wdm_thermal(const aar_params* p) { return p->WDM_On == 1 && p->WDM_Vtherm_On == 1; }
}  // namespace

// ------------------------------------------------------------------ writers
static int fwrite_all(const void* p, size_t sz, size_t n, FILE* fd) {
  return fwrite(p, sz, n, fd) == n ? AAR_OK : 777;
}

// Gadget-1 header of save_local_data (save.c:195-266) for n of ntot particles in numfiles files.
static void gadget_header(uint8_t* hdr, const aar_params* p, const aar_cosmo* c, const uint8_t* glass_header,
                          long long n, long long ntot, int numfiles) {
  if (glass_header) memcpy(hdr, glass_header, 256);  // fields not set below keep the glass's bytes
  else memset(hdr, 0, 256);
  unsigned npart6[6] = {0, 0, 0, 0, 0, 0}, ntot6[6] = {0, 0, 0, 0, 0, 0};
  double mass6[6] = {0, 0, 0, 0, 0, 0};
  npart6[1] = (unsigned)n;
  ntot6[1] = (unsigned)ntot;
  ntot6[2] = (unsigned)(ntot >> 32);
  mass6[1] = (p->Omega) * 3 * c->Hubble * c->Hubble / (8 * kPi * c->G) * pow(p->Box, 3) / ntot;
  double time = c->InitTime, redshift = 1.0 / c->InitTime - 1;
  int zero = 0;
  memcpy(hdr + 0, npart6, 24);
  memcpy(hdr + 24, mass6, 48);
  memcpy(hdr + 72, &time, 8);
  memcpy(hdr + 80, &redshift, 8);
  memcpy(hdr + 88, &zero, 4);   // flag_sfr
  memcpy(hdr + 92, &zero, 4);   // flag_feedback
  memcpy(hdr + 96, ntot6, 24);
  memcpy(hdr + 120, &zero, 4);  // flag_cooling
  memcpy(hdr + 124, &numfiles, 4);
  memcpy(hdr + 128, &p->Box, 8);
  memcpy(hdr + 136, &p->Omega, 8);
  memcpy(hdr + 144, &p->OmegaLambda, 8);
  memcpy(hdr + 152, &p->HubbleParam, 8);
  memcpy(hdr + 160, &zero, 4);  // flag_stellarage
  memcpy(hdr + 164, &zero, 4);  // flag_metals
  memcpy(hdr + 168, &zero, 4);  // hashtabsize
}

// One Gadget-1 file of save_local_data (save.c:162-520, NO64BITID): n particles
// starting at global index first, of ntot in numfiles files.
// fdist (WDM thermal speeds, or NULL): the file's writer draws them from drand48's first state on.
static int write_gadget_file(const std::string& fn, const aar_params* p, const aar_cosmo* c, const float* pos,
                             const float* vel, long long first, long long n, long long ntot, int numfiles,
                             const uint8_t* glass_header, const FermiDirac* fdist) {
  FILE* fd = fopen(fn.c_str(), "wb");
  if (!fd) return 10;
  uint8_t hdr[256];
  gadget_header(hdr, p, c, glass_header, n, ntot, numfiles);
  int rc = AAR_OK;
  int dummy = 256;
  rc |= fwrite_all(&dummy, 4, 1, fd);
  rc |= fwrite_all(hdr, 256, 1, fd);
  rc |= fwrite_all(&dummy, 4, 1, fd);
  // record markers are 4-byte ints, truncated as in the reference for > 2^31 bytes
  dummy = (int)(uint32_t)(sizeof(float) * 3 * (size_t)n);
  rc |= fwrite_all(&dummy, 4, 1, fd);
  rc |= fwrite_all(pos + 3 * first, sizeof(float), 3 * (size_t)n, fd);
  rc |= fwrite_all(&dummy, 4, 1, fd);
  rc |= fwrite_all(&dummy, 4, 1, fd);
  if (fdist) {
    Drand48 rng;
    std::vector<float> vb(3 << 20);
    for (long long s = 0; s < n; s += (long long)vb.size() / 3) {
      const long long m = std::min<long long>((long long)vb.size() / 3, n - s);
      memcpy(vb.data(), vel + 3 * (first + s), sizeof(float) * 3 * (size_t)m);
      for (long long q = 0; q < m; ++q) fdist->kick(&vb[3 * q], rng);
      rc |= fwrite_all(vb.data(), sizeof(float), 3 * (size_t)m, fd);
    }
  } else {
    rc |= fwrite_all(vel + 3 * first, sizeof(float), 3 * (size_t)n, fd);
  }
  rc |= fwrite_all(&dummy, 4, 1, fd);
  dummy = (int)(uint32_t)(sizeof(int) * (size_t)n);
  rc |= fwrite_all(&dummy, 4, 1, fd);
  std::vector<int> ids(1 << 20);
  for (long long s = 0; s < n; s += (long long)ids.size()) {
    const long long m = std::min<long long>((long long)ids.size(), n - s);
    // IDs count from 1 over the whole run (read_glass.c:176)
    for (long long q = 0; q < m; ++q) ids[q] = (int)(first + s + q + 1);
    rc |= fwrite_all(ids.data(), sizeof(int), (size_t)m, fd);
  }
  rc |= fwrite_all(&dummy, 4, 1, fd);
  if (fclose(fd)) rc = 777;
  return rc ? 777 : AAR_OK;
}

// Gadget-1 snapshot of one task (save.c:162-520, NO64BITID).
extern "C" int aar_write_gadget(const aar_params* p, const aar_cosmo* c, const float* pos,
                                const float* vel, long long npart, const uint8_t* glass_header) {
  if (!p || !c || !pos || !vel || npart < 0) return AAR_ERR_BAD_ARG;
  if (p->NumFilesWrittenInParallel > 1) return AAR_ERR_WRITE_PARALLEL;  // save.c:112-117, NTask = 1
  if (npart == 0) return AAR_OK;
  std::unique_ptr<FermiDirac> fdist(wdm_thermal(p) ? new FermiDirac(p) : nullptr);
  return write_gadget_file(std::string(p->OutputDir) + "/" + p->FileBase, p, c, pos, vel, 0, npart, npart, 1,
                           glass_header, fdist.get());
}

// The snapshot split over files the way several tasks write it (save.c:184-187, 252:
// OutputDir/FileBase.<i>, num_files = the file count, npart per file,
// npartTotal of the run), so that every record stays below max_record bytes.
extern "C" int aar_write_gadget_files(const aar_params* p, const aar_cosmo* c, const float* pos,
                                      const float* vel, long long npart, const uint8_t* glass_header, int nfiles,
                                      long long max_record, int* nfiles_out) {
  if (!p || !c || !pos || !vel || npart < 0 || nfiles < 0 || max_record < 0) return AAR_ERR_BAD_ARG;
  if (npart == 0) {  // save.c:181-182: a task without particles writes no file
    if (nfiles_out) *nfiles_out = 0;
    return AAR_OK;
  }
  if (max_record == 0 || max_record > 2147483647LL) max_record = 2147483647LL;
  const long long per_max = max_record / (3 * (long long)sizeof(float));  // position block is the largest
  if (per_max < 1) return AAR_ERR_BAD_ARG;
  if (nfiles == 0) {
    nfiles = (int)std::max<long long>(1, (npart + per_max - 1) / per_max);
    while ((npart + nfiles - 1) / nfiles > per_max) ++nfiles;
  }
  if ((npart + nfiles - 1) / nfiles > per_max || nfiles > npart) return AAR_ERR_BAD_ARG;
  if (nfiles_out) *nfiles_out = nfiles;
  const std::string base = std::string(p->OutputDir) + "/" + p->FileBase;
  std::unique_ptr<FermiDirac> fdist(wdm_thermal(p) ? new FermiDirac(p) : nullptr);
  for (int f = 0; f < nfiles; ++f) {  // each file as its own writing task
    const long long a = npart * f / nfiles, b = npart * (f + 1) / nfiles;
    const std::string fn = nfiles > 1 ? base + "." + std::to_string(f) : base;
    const int rc = write_gadget_file(fn, p, c, pos, vel, a, b - a, npart, nfiles, glass_header, fdist.get());
    if (rc) return rc;
  }
  return AAR_OK;
}

// x-planes of task r of ntask in FFTW 2's MPI slab decomposition
// (rfftwnd_mpi_local_sizes, main.c:1573-1588): blocks of ceil(N / ntask)
// planes, the last non-empty task takes the remainder, later tasks none.
static void fftw_slab(int N, int ntask, int r, int* start, int* nx) {
  const int block = (N + ntask - 1) / ntask, used = (N + block - 1) / block;
  if (r >= used) {
    *start = 0;
    *nx = 0;
  } else {
    *start = r * block;
    *nx = r == used - 1 ? N - *start : block;
  }
}

// Potential records of planes [x0, x0 + nx) of a real N^3 field (save.c:88-99).
static int write_phi_file(const char* name, const aar_params* p, const double* real, int x0, int nx, int total) {
  const int N = p->Nmesh, W = 2 * (N / 2 + 1);
  FILE* fd = fopen(name, "wb");
  if (!fd) return 10;
  int rc = fwrite_all(&total, 4, 1, fd);
  rc |= fwrite_all(&N, 4, 1, fd);
  rc |= fwrite_all(&p->Box, 8, 1, fd);
  // the reference issues two fwrite calls per element; here one x-plane of
  // {int coord, double value} records is packed and written at a time
  std::vector<unsigned char> rec((size_t)N * N * 12);
  for (int i = 0; i < nx && !rc; ++i) {
    unsigned char* o = rec.data();
    for (int j = 0; j < N; ++j) {
      const double* src = real + ((size_t)(x0 + i) * N + j) * N;
      for (int k = 0; k < N; ++k, o += 12) {
        const int coord = (i * N + j) * W + k;  // slab-local, as each task stores it
        memcpy(o, &coord, 4);
        memcpy(o + 4, &src[k], 8);
      }
    }
    rc |= fwrite_all(rec.data(), 1, rec.size(), fd);
  }
  if (fclose(fd)) rc = 777;
  return rc ? 777 : AAR_OK;
}

// Real-space potential dump (save.c:14-105), one task.
extern "C" int aar_write_phi(const aar_params* p, const double* real, int non_gaussian) {
  if (!p || !real) return AAR_ERR_BAD_ARG;
  const int N = p->Nmesh, W = 2 * (N / 2 + 1);
  char name[1200];
  snprintf(name, sizeof(name), "%s/%s_potential_%d", p->OutputDir, non_gaussian ? "Nongauss" : "Gauss",
           p->Seed);
  return write_phi_file(name, p, real, 0, N, N * N * W + N * W);  // TotalSizePlusAdditional of one task
}

// The dump as ntask tasks write it (save.c:14-105): one file per task with its
// slab, named ..._StartX_%04d_NumX_%d when more than one task holds particles;
// every header carries task 0's TotalSizePlusAdditional.
extern "C" int aar_write_phi_tasks(const aar_params* p, const double* real, int non_gaussian, int ntask,
                                   int ntask_with_n) {
  if (!p || !real || ntask < 1) return AAR_ERR_BAD_ARG;
  if (ntask_with_n <= 1) return aar_write_phi(p, real, non_gaussian);
  const int N = p->Nmesh, W = 2 * (N / 2 + 1);
  int s0, n0;
  fftw_slab(N, ntask, 0, &s0, &n0);
  const int total = n0 * N * W + N * W;
  for (int r = 0; r < ntask; ++r) {
    int x0, nx;
    fftw_slab(N, ntask, r, &x0, &nx);
    char name[1200];
    snprintf(name, sizeof(name), "%s/%s_potential_%d_StartX_%04d_NumX_%d", p->OutputDir,
             non_gaussian ? "Nongauss" : "Gauss", p->Seed, x0, nx);
    const int rc = write_phi_file(name, p, real, x0, nx, total);
    if (rc) return rc;
  }
  return AAR_OK;
}

// Task of every glass particle in every x-tile: the slab of its unperturbed x
// (read_glass.c:129-146, x stored as float), task_of[ti * nglass + n].
static void glass_tasks(const aar_params* p, const aar_tables* t, int ntask, std::vector<uint16_t>& task_of) {
  const int N = p->Nmesh, T = p->GlassTileFac;
  std::vector<int> slab_task(N);
  for (int r = 0; r < ntask; ++r) {
    int x0, nx;
    fftw_slab(N, ntask, r, &x0, &nx);
    for (int i = 0; i < nx; ++i) slab_task[x0 + i] = r;
  }
  task_of.assign((size_t)T * t->nglass, 0);
  for (int ti = 0; ti < T; ++ti)
    for (int n = 0; n < t->nglass; ++n) {
      const float x = (float)(t->glass[3 * (size_t)n] / t->glass_box * (p->Box / T) + ti * (p->Box / T));
      // the reference truncates and clamps at N - 1; a position outside the box (or
      // NaN) would index outside its slab table, here it goes to the nearest end
      const double sd = (double)x / p->Box * N;
      const int slab = !(sd >= 0) ? 0 : (sd >= N ? N - 1 : (int)sd);
      task_of[(size_t)ti * t->nglass + n] = (uint16_t)slab_task[slab];
    }
}

extern "C" int aar_task_layout(const aar_params* p, const aar_tables* t, int ntask, long long* npart_task,
                               int* ntask_with_n) {
  if (!p || !t || !t->glass || t->nglass < 1 || ntask < 1 || ntask > 65535 || p->GlassTileFac < 1 || p->Nmesh < 1)
    return AAR_ERR_BAD_ARG;
  std::vector<uint16_t> task_of;
  glass_tasks(p, t, ntask, task_of);
  std::vector<long long> cnt(ntask, 0);
  const long long T2 = (long long)p->GlassTileFac * p->GlassTileFac;
  for (size_t q = 0; q < task_of.size(); ++q) cnt[task_of[q]] += T2;  // every (tj, tk) repeats the x-tile
  int with = 0;
  for (int r = 0; r < ntask; ++r) {
    if (npart_task) npart_task[r] = cnt[r];
    with += cnt[r] > 0;
  }
  if (ntask_with_n) *ntask_with_n = with;
  return AAR_OK;
}

// The snapshot as ntask tasks write it (write_particle_data, save.c:108-520):
// task r's particles -- those whose unperturbed x lies in its slabs -- in
// file order, with their run-wide IDs, in OutputDir/FileBase.r (FileBase when
// one task holds them all); tasks without particles write no file.
extern "C" int aar_write_gadget_tasks(const aar_params* p, const aar_cosmo* c, const aar_tables* t,
                                      const float* pos, const float* vel, long long npart, int ntask) {
  if (!p || !c || !t || !pos || !vel || !t->glass || ntask < 1 || ntask > 65535) return AAR_ERR_BAD_ARG;
  if (p->NumFilesWrittenInParallel > ntask) return AAR_ERR_WRITE_PARALLEL;  // save.c:112-117
  const int T = p->GlassTileFac;
  const long long ng = t->nglass;
  if ((long long)T * T * T * ng != npart) return AAR_ERR_BAD_ARG;
  std::vector<uint16_t> task_of;
  glass_tasks(p, t, ntask, task_of);
  std::vector<long long> cnt(ntask, 0);
  for (size_t q = 0; q < task_of.size(); ++q) cnt[task_of[q]] += (long long)T * T;
  int with = 0;
  for (int r = 0; r < ntask; ++r) with += cnt[r] > 0;
  std::unique_ptr<FermiDirac> fdist(wdm_thermal(p) ? new FermiDirac(p) : nullptr);  // task 0 prints
  if (with <= 1) {
    for (int r = 0; r < ntask; ++r)
      if (cnt[r] == npart)
        return write_gadget_file(std::string(p->OutputDir) + "/" + p->FileBase, p, c, pos, vel, 0, npart, npart,
                                 1, t->glass_header, fdist.get());
    return AAR_OK;
  }
  // x-tiles holding each task's particles
  std::vector<int> tlo(T, 65535), thi(T, -1);
  for (int ti = 0; ti < T; ++ti)
    for (long long n = 0; n < ng; ++n) {
      tlo[ti] = std::min<int>(tlo[ti], task_of[(size_t)ti * ng + n]);
      thi[ti] = std::max<int>(thi[ti], task_of[(size_t)ti * ng + n]);
    }
  std::vector<float> buf(3 << 20);
  std::vector<int> ids(1 << 20);
  for (int r = 0; r < ntask; ++r) {
    if (!cnt[r]) continue;
    const std::string fn = std::string(p->OutputDir) + "/" + p->FileBase + "." + std::to_string(r);
    FILE* fd = fopen(fn.c_str(), "wb");
    if (!fd) return 10;
    uint8_t hdr[256];
    gadget_header(hdr, p, c, t->glass_header, cnt[r], npart, with);
    int rc = AAR_OK, dummy = 256;
    rc |= fwrite_all(&dummy, 4, 1, fd);
    rc |= fwrite_all(hdr, 256, 1, fd);
    rc |= fwrite_all(&dummy, 4, 1, fd);
    // pass 0: positions, 1: velocities (with the task's own thermal speeds), 2: IDs, each gathered
    // in particle order
    Drand48 rng;  // every task's drand48 starts from the default state
    for (int pass = 0; pass < 3 && !rc; ++pass) {
      dummy = (int)(uint32_t)((pass < 2 ? 12 : 4) * (size_t)cnt[r]);
      rc |= fwrite_all(&dummy, 4, 1, fd);
      size_t fill = 0;
      auto flush = [&]() {
        if (pass < 2) rc |= fwrite_all(buf.data(), sizeof(float), 3 * fill, fd);
        else rc |= fwrite_all(ids.data(), sizeof(int), fill, fd);
        fill = 0;
      };
      for (int ti = 0; ti < T; ++ti) {
        if (r < tlo[ti] || r > thi[ti]) continue;
        const uint16_t* to = &task_of[(size_t)ti * ng];
        for (int tj = 0; tj < T; ++tj)
          for (int tk = 0; tk < T; ++tk) {
            const long long base = (((long long)ti * T + tj) * T + tk) * ng;
            for (long long n = 0; n < ng; ++n) {
              if (to[n] != r) continue;
              const long long q = base + n;
              if (pass == 2) {
                ids[fill] = (int)(q + 1);
              } else {
                const float* src = (pass == 0 ? pos : vel) + 3 * q;
                buf[3 * fill] = src[0];
                buf[3 * fill + 1] = src[1];
                buf[3 * fill + 2] = src[2];
                if (pass == 1 && fdist) fdist->kick(&buf[3 * fill], rng);
              }
              if (++fill == ids.size()) flush();
            }
          }
      }
      flush();
      rc |= fwrite_all(&dummy, 4, 1, fd);
    }
    if (fclose(fd)) rc = 777;
    if (rc) return 777;
  }
  return AAR_OK;
}

// inputspec_<FileBase>.txt with the Peacock & Dodds nonlinear spectrum (main.c:1620-1694).
extern "C" int aar_write_inputspec(const aar_params* p, const aar_tables* t, const aar_cosmo* c) {
  if (!p || !c) return AAR_ERR_BAD_ARG;
  Cosmo cm(*p, t);
  if (cm.status) return cm.status;
  cm.norm = c->Norm;
  char name[1200];
  snprintf(name, sizeof(name), "%s/inputspec_%s.txt", p->OutputDir, p->FileBase);
  FILE* fd = fopen(name, "w");
  if (!fd) return 10;
  double gf = cm.growth_factor(0.001, 1.0) / (1.0 / 0.001);
  double DDD = cm.growth_factor(1.0 / (p->Redshift + 1), 1.0);
  fprintf(fd, "%12g %12g\n", p->Redshift, DDD);
  double kstart = 2 * kPi / (1000.0 * (3.085678e24 / p->UnitLength_in_cm));
  double kend = 2 * kPi / (0.001 * (3.085678e24 / p->UnitLength_in_cm));
  for (double k = kstart; k < kend; k *= 1.025) {
    double po = cm.powerspec(k);
    double dl = 4.0 * kPi * k * k * k * po;
    double kf = 0.5;
    double po2 = cm.powerspec(1.001 * k * kf);
    double po1 = cm.powerspec(k * kf);
    double dnl = 0, knl = 0;
    if (po != 0 && po1 != 0 && po2 != 0) {
      double neff = (log(po2) - log(po1)) / (log(1.001 * k * kf) - log(k * kf));
      if (1 + neff / 3 > 0) {
        double A = 0.482 * pow(1 + neff / 3, -0.947);
        double B = 0.226 * pow(1 + neff / 3, -1.778);
        double alpha = 3.310 * pow(1 + neff / 3, -0.244);
        double beta = 0.862 * pow(1 + neff / 3, -0.287);
        double V = 11.55 * pow(1 + neff / 3, -0.423) * 1.2;
        double x = dl;
        dnl = x * pow((1 + B * beta * x + pow(A * x, alpha * beta)) /
                          (1 + pow(pow(A * x, alpha) * gf * gf * gf / (V * sqrt(x)), beta)),
                      1 / beta);
        knl = k * pow(1 + dnl, 1.0 / 3);
      }
    }
    fprintf(fd, "%12g %12g    %12g %12g\n", k, dl, knl, dnl);
  }
  fclose(fd);
  return AAR_OK;
}
