// 2LPTBasis <ParameterFile> [--fp32] [--plain] [--device N] [--max-record BYTES]
// 2LPTResBasis <ParameterFile> [...]   (this file built with -DAAR_RESBASIS)
//
// Drop-in for the reference executables Aarambam/LPT/Basis and
// Aarambam/LPT/ResBasis (main.c:53-102) running as a single task on one GPU:
// same parameter file, tables, glass, stdout stage banners, Gadget snapshot,
// optional potential dumps and inputspec file.  The default is the FP64
// pairwise schedule, the closest to the reference; --fp32 selects the FP32
// batched schedule.  The snapshot is one file unless a record would exceed
// 2^31 - 1 bytes (or --max-record BYTES): then FileBase.<i> files.  Under
// mpirun / srun with N tasks, task 0 runs the job and writes the outputs in
// the layout of N reference tasks; the other tasks exit.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <vector>

#include "../../include/aarambam.h"

static double now() {
  timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return t.tv_sec + 1e-9 * t.tv_nsec;
}

static void die(int code, const char* what) {
  printf("FatalError called with number=%d (%s: %s)\n", code, what, aar_status_string(code));
  fflush(stdout);
  exit(code > 255 ? 1 : (code ? code : 1));
}

#ifdef AAR_RESBASIS
static const bool kRes = true;
static const char kExec[] = "Aarambam-2LPT-ResBasis";
#else
static const bool kRes = false;
static const char kExec[] = "Aarambam-2LPT-Basis";
#endif

static void banner(const aar_params& p, int nglass) {  // main.c:104-139 (ResBasis main.c:104-137)
  char pstr[80];
  memset(pstr, '*', 79);
  pstr[79] = 0;
  printf("%s\n", pstr);
  // the reference's title line ends at the terminator its sprintf leaves after the name
  printf("*%*s%s\n", 28, "", kExec);
  printf("%s\n\n", pstr);
  printf("      Box = %.2e   Nmesh = %d   Nsample = %d   Nmax = %d\n", p.Box, p.Nmesh, p.Nsample, p.Nmax);
  printf("                       Nglass = %d    GlassTileFac = %d\n\n", nglass, p.GlassTileFac);
  printf("        Omega = %.4f         OmegaLambda = %.4f    OmegaBaryon = %.2e\n", p.Omega, p.OmegaLambda,
         p.OmegaBaryon);
  printf("       sigma8 = %.4f     PrimoridalIndex = %.4f       Redshift = %.2e\n", p.Sigma8,
         p.PrimordialIndex, p.Redshift);
  if (kRes) {
    printf("       A_Pk = %.4f     w_Pk = %.4f       phi_Pk = %.4f\n", p.A_res, p.w_res, p.phi_res);
    printf("  HubbleParam = %.4f  OmegaDM_2ndSpecies = %.2e          fNL = %+.2e\n\n", p.HubbleParam,
           p.OmegaDM_2ndSpecies, p.Fnl);
  } else {
    printf("   HubbleParam = %.4f  OmegaDM_2ndSpecies = %.2e          fNL = %+.2e\n\n", p.HubbleParam,
           p.OmegaDM_2ndSpecies, p.Fnl);
  }
  printf("   FixedAmplitude = %d    PhaseFlip = % d   SphereMode = %d    Seed = %d\n", p.FixedAmplitude,
         p.PhaseFlip, p.SphereMode, p.Seed);
  printf("\n%s\n\n", pstr);
}

// Tasks of an MPI launch (mpirun / srun set these before the program starts);
// 1 and rank 0 when run directly.
static void mpi_world(int* size, int* rank) {
  const char* sz[] = {"OMPI_COMM_WORLD_SIZE", "PMI_SIZE", "MV2_COMM_WORLD_SIZE", "SLURM_NTASKS"};
  const char* rk[] = {"OMPI_COMM_WORLD_RANK", "PMI_RANK", "MV2_COMM_WORLD_RANK", "SLURM_PROCID"};
  *size = 1;
  *rank = 0;
  for (int q = 0; q < 4; ++q) {
    const char* a = getenv(sz[q]);
    const char* b = getenv(rk[q]);
    if (a && b && atoi(a) > 0) {
      *size = atoi(a);
      *rank = atoi(b);
      return;
    }
  }
}

int main(int argc, char** argv) {
  // Launched as `mpirun -np N` (as the reference's examples do): task 0 runs the
  // whole job on one GPU and writes the outputs in the layout N tasks write;
  // the other tasks have nothing to do.
  int ntask, rank;
  mpi_world(&ntask, &rank);
  if (rank != 0) return 0;
  if (argc < 2) {  // main.c:65-72
    fprintf(stdout, "\nParameters are missing.\n");
    fprintf(stdout, "Call with <ParameterFile>\n\n");
    return 0;
  }
  int precision = AAR_FP64, schedule = AAR_SCHED_PLAIN, device = 0;
  long long max_record = 2147483647LL;
  for (int a = 2; a < argc; ++a) {
    if (!strcmp(argv[a], "--fp32")) {
      precision = AAR_FP32;
      schedule = AAR_SCHED_BATCHED;
    } else if (!strcmp(argv[a], "--batched")) {
      schedule = AAR_SCHED_BATCHED;
    } else if (!strcmp(argv[a], "--plain")) {
      schedule = AAR_SCHED_PLAIN;
    } else if (!strcmp(argv[a], "--device") && a + 1 < argc) {
      device = atoi(argv[++a]);
    } else if (!strcmp(argv[a], "--max-record") && a + 1 < argc) {
      max_record = atoll(argv[++a]);
      if (max_record < 12 || max_record > 2147483647LL) {
        printf("--max-record takes 12 to 2147483647 bytes\n");
        return 1;
      }
    }
  }
  const double t0 = now();
  printf("RUNNING LPTPNG WITH %d TASKS", ntask);
  aar_params p;
  char msg[8192];
  const int prc = kRes ? aar_read_params_res(argv[1], &p, msg, sizeof(msg))
                       : aar_read_params(argv[1], &p, msg, sizeof(msg));
  if (prc) {  // read_param.c:241-255: report, exit(0)
    fputs(msg, stdout);
    return 0;
  }
  if (p.NumFilesWrittenInParallel > ntask) die(24131, "NumFilesWrittenInParallel");  // save.c:112-117
  std::vector<double> alpha((size_t)p.N_modes * p.N_modes * p.N_modes), x0(2 * (size_t)p.N_modes);
  int rc;
  if ((rc = aar_load_alpha(p.FileWithAlphaCoeff, p.N_modes, alpha.data()))) die(rc, p.FileWithAlphaCoeff);
  if ((rc = aar_load_x0(p.FileWithX0Coeff, p.N_modes, x0.data()))) die(rc, p.FileWithX0Coeff);
  std::vector<double> tk, tt;
  aar_tables t{};
  if (p.WhichTransfer == 2) {
    int n = 0;
    if ((rc = aar_load_transfer(p.FileWithInputTransfer, &n, nullptr, nullptr)))
      die(rc, p.FileWithInputTransfer);
    tk.resize(n);
    tt.resize(n);
    if ((rc = aar_load_transfer(p.FileWithInputTransfer, &n, tk.data(), tt.data())))
      die(rc, p.FileWithInputTransfer);
    t.ntransfer = n;
    t.transfer_k = tk.data();
    t.transfer_t = tt.data();
  }
  int nglass = 0;
  double gbox = 0;
  uint8_t ghdr[256];
  if ((rc = aar_read_glass(p.GlassFile, &nglass, nullptr, &gbox, ghdr))) die(rc, p.GlassFile);
  std::vector<float> glass(3 * (size_t)nglass);
  if ((rc = aar_read_glass(p.GlassFile, &nglass, glass.data(), &gbox, ghdr))) die(rc, p.GlassFile);
  t.alpha = alpha.data();
  t.x0 = x0.data();
  t.glass = glass.data();
  t.nglass = nglass;
  t.glass_box = gbox;
  t.glass_header = ghdr;
  aar_cosmo c;
  if ((rc = aar_cosmology(&p, &t, &c))) die(rc, "cosmology");
  printf("\nNormalization of spectrum in file:  Sigma8 = %g\n", c.sigma8_in);
  printf("Normalization adjusted to  Sigma8=%g   (Normfac=%g)\n\n", p.Sigma8, c.Norm);
  banner(p, nglass);
  aar_ctx* ctx = nullptr;
  if ((rc = aar_ctx_create(device, &ctx))) die(rc, "aar_ctx_create");
  aar_job* job = nullptr;
  const double tp = now();
  if ((rc = aar_job_prepare(ctx, &p, &t, precision, schedule, &job))) die(rc, "aar_job_prepare");
  const double tr = now();
  printf("Computing non-Gaussian potential and 2LPT displacements on the GPU (%s, %s)...",
         precision == AAR_FP32 ? "fp32" : "fp64", schedule == AAR_SCHED_BATCHED ? "batched" : "plain");
  fflush(stdout);
  if ((rc = aar_job_run(job, nullptr)) || (rc = aar_job_sync(job))) die(rc, "aar_job_run");
  printf(" Done [%.3f s on device, %.3f s prepare]\n", aar_job_last_seconds(job), tr - tp);
  const long long npart = (long long)p.GlassTileFac * p.GlassTileFac * p.GlassTileFac * nglass;
  int with_n = 1;
  if (ntask > 1 && (rc = aar_task_layout(&p, &t, ntask, nullptr, &with_n))) die(rc, "task layout");
  if (p.SavePotentialField == 1) {  // main.c:1096, 1135
    std::vector<double> f((size_t)p.Nmesh * p.Nmesh * p.Nmesh);
    if ((rc = aar_job_fetch_field(job, AAR_FIELD_PHIG_REAL, f.data())) ||
        (rc = aar_write_phi_tasks(&p, f.data(), 0, ntask, with_n)))
      die(rc, "write_phi");
    if ((rc = aar_job_fetch_field(job, AAR_FIELD_PHING_REAL, f.data())) ||
        (rc = aar_write_phi_tasks(&p, f.data(), 1, ntask, with_n)))
      die(rc, "write_phi");
  }
  std::vector<float> pos(3 * (size_t)npart), vel(3 * (size_t)npart);
  if ((rc = aar_job_fetch(job, pos.data(), vel.data()))) die(rc, "fetch");
  printf("Writing initial conditions snapshot...");
  fflush(stdout);
  const double tw = now();
  // One file as the reference writes it, unless its position block would overflow Gadget-1's
  // 4-byte record markers (12 npart >= 2^31, e.g. 1024^3) or --max-record: then FileBase.<i>
  // files as several tasks write them.
  int nfiles = 1;
  if (ntask > 1) {
    if ((rc = aar_write_gadget_tasks(&p, &c, &t, pos.data(), vel.data(), npart, ntask))) die(rc, "write_particle_data");
    nfiles = with_n;
  } else if (12LL * npart <= max_record) {
    if ((rc = aar_write_gadget(&p, &c, pos.data(), vel.data(), npart, ghdr))) die(rc, "write_particle_data");
  } else if ((rc = aar_write_gadget_files(&p, &c, pos.data(), vel.data(), npart, ghdr, 0, max_record,
                                          &nfiles))) {
    die(rc, "write_particle_data");
  }
  printf(" Done [%.3f s, %d file%s]\n", now() - tw, nfiles, nfiles > 1 ? "s" : "");
  aar_job_destroy(job);
  aar_ctx_destroy(ctx);
  if ((rc = aar_write_inputspec(&p, &t, &c))) die(rc, "print_spec");
  printf("Total wall time %.3f s\n", now() - t0);
  return 0;
}
