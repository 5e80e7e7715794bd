// Job orchestration and the device half of the C-ABI.
//
// A job owns every device buffer of one initial-condition run: a pool of
// spectrum-sized field slots, the seed table, the k-space tables, the glass
// and the particle arrays.  aar_job_run only enqueues kernels on a stream,
// so several jobs can be in flight on separate streams.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <vector>

#include "../../include/aarambam.h"
#include "aar_common.h"
#include "aar_internal.h"
#include "aar_jit.h"
#include "aar_kernels.h"

int aar_host_transfer_tables(const aar_params* p, const aar_tables* t, std::vector<double>& logk,
                             std::vector<double>& logT);

namespace {

struct PairWork {  // one (N_i, N_j) product of the pairwise schedule
  int i, j;
  int w_off, w_n;  // slice of the weight list
};

template <typename T> struct Exec;

// The device-resident tables of a job: everything its kernels read except the seed table,
// fNL and the job's own buffers (field slots, particles).  One plan serves every job of the
// same parameters (Seed and Fnl aside), tables, precision and schedule: the context keeps
// plans in a cache, so a stream of jobs that differ in seed or fNL uploads its tables once.
// The device block is one buffer holding every table at 256-byte aligned offsets.
struct Plan {
  std::string key;
  std::vector<double> alpha, x0in, tk, tt;  // the caller's tables, matched exactly on lookup
  std::vector<float> glass_in;
  aar_cosmo c{};
  int M = 0, K = 0;
  std::vector<int> modes_h, kmodes_h, groups_h;
  std::vector<NgTerm> terms_h;
  std::vector<PairWork> pairs;
  std::vector<int> fgroups_h;  // the factored product schedule (host copy, for aar_jit.cpp)
  std::vector<NgRec4> frecs_h;
  std::vector<char> image;  // host image of the block
  size_t bytes = 0;
  void* block = nullptr;
  bool owned = false;  // hipMalloc'd by the cache (freed with the plan)
  size_t o_kv = 0, o_sinc = 0, o_icic = 0, o_x0 = 0, o_outc = 0, o_modes = 0, o_kmodes = 0, o_groups = 0, o_terms = 0,
         o_weights = 0, o_fgroups = 0, o_frecs = 0, o_logk = 0, o_logT = 0,
         o_glass = 0;
  int nglass = 0;
  double glass_box = 0;
  int nt = 0;  // transfer table rows (WhichTransfer 2)
  unsigned long long last_use = 0;
  ~Plan() {
    if (owned && block) hipFree(block);
  }
};

}  // namespace

struct aar_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // batched jobs run the NG sum's fused z pass compiled for their own product schedule
  // (aar_jit.cpp) unless turned off (aar_ctx_set_jit)
  bool jit = true;
  // batched jobs filter the mode fields in the x pass's loads (Pipeline::modes_xy): 0 never,
  // 1 when the filters' Legendre recurrences are short (aar_ctx_set_modes_x), 2 always
  int modes_x = 1;
  float2* tw32 = nullptr;
  double2* tw64 = nullptr;
  // cosmology scalars by the parameters they depend on: the Romberg sigma8
  // normalisation and growth-factor integrals are the costliest host step of
  // a job and repeat across jobs that differ only in mesh, table or seed
  std::mutex mu;
  std::map<std::string, aar_cosmo> cosmo_cache;
  // device-resident plans (tables) by key, least recently used evicted first
  std::mutex plan_mu;
  std::vector<std::shared_ptr<Plan>> plans;
  unsigned long long plan_clock = 0;
};

struct aar_job {
  aar_ctx* ctx = nullptr;
  aar_params p{};
  aar_cosmo c{};
  int prec = AAR_FP64, sched = AAR_SCHED_BATCHED;
  Geom g{};
  size_t cxsize = 16;
  // device buffers
  void* pool = nullptr;
  int nslots = 0;
  uint32_t* seeds = nullptr;
  double *kv = nullptr, *sinc = nullptr, *icic = nullptr, *x0 = nullptr, *logk = nullptr, *logT = nullptr;
  unsigned char* outc = nullptr;
  int *modes = nullptr, *kmodes = nullptr, *groups = nullptr;
  NgTerm* terms = nullptr;
  int* fgroups = nullptr;  // the terms factored for the fused z pass
  NgRec4* frecs = nullptr;
  NgWeight* weights = nullptr;
  double2* acc = nullptr;
  float* glass = nullptr;
  float *pos = nullptr, *vel = nullptr;
  long long npart = 0;
  int nglass = 0;
  double glass_box = 0;
  KSpace ks{};
  TransferDev tf{};
  // schedule
  int M = 0, K = 0;
  std::vector<int> modes_h, kmodes_h;
  std::vector<NgTerm> terms_h;  // regrouped terms (mode slots i, j; coefficient) by outer index
  std::vector<int> groups_h;
  std::vector<PairWork> pairs;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_seconds = 0;
  // per-kind launch timing
  bool profiling = false;
  struct Rec {
    int kind;
    double bytes;
    hipEvent_t a, b;
    int dispatches;  // kernel launches between the two events
  };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> evpool;
  size_t evnext = 0;
  int k_launch[AAR_NKINDS] = {};
  double k_sec[AAR_NKINDS] = {}, k_bytes[AAR_NKINDS] = {};
  hipStream_t last_stream = nullptr;
  bool ran = false;
  // hipGraph of the whole job (captured on the first run with graph mode on)
  bool use_graph = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  // device memory carved from a caller's workspace (serving arena) instead of
  // hipMalloc; host copies of uploaded tables live until destroy when the
  // uploads are asynchronous on `up`
  char* ws = nullptr;
  size_t ws_size = 0, ws_used = 0;
  hipStream_t up = nullptr;
  std::vector<std::vector<char>> staged;
  // the tables this job reads (shared through the context's cache, or its own in a workspace)
  std::shared_ptr<Plan> plan;
  bool ext_out = false;  // pos / vel are the caller's device arrays
  // seed tables of later runs (aar_job_reseed): built on the host into one of two pinned
  // buffers and copied to `seeds` at the start of the next run; an event after each copy
  // tells when its buffer may be rewritten
  uint32_t* stage[2] = {nullptr, nullptr};
  hipEvent_t stage_ev[2] = {nullptr, nullptr};
  bool stage_used[2] = {false, false};
  int stage_next = 0, stage_pending = -1;
  bool fused_lpt = false;
  // the mode fields' filters need few Legendre steps per value (modes_xy pays off, at 1024^3 L+E;
  // at 512^3 with 16 modes, 102 steps per value over the fields, the separate filter pass is
  // faster since it advances one recurrence through all fields)
  bool short_filters = false;
  // the NG sum's fused z pass compiled for this job's schedule (jit_status: the status of a
  // failed compile, when the library's k_z_ng runs instead)
  JitZng jit;
  int jit_status = 0;
  // slot map: phi_G, Phi, the work region; ZA and 2LPT displacement fields
  int s_phig = 0, s_phi = 1, s_reg = 2, s_disp = 8, s_disp2 = 2;
  void* slot(int s) const { return (char*)pool + (size_t)s * g.fs * cxsize; }
};

namespace {

bool pow2(int n) { return n >= 16 && n <= 1024 && (n & (n - 1)) == 0; }
int ilog2i(int n) {
  int l = 0;
  while ((1 << l) < n) ++l;
  return l;
}

struct NgFactored {
  std::vector<int> groups;  // K + 1 offsets into recs, in records
  std::vector<NgRec4> recs;
};

// Build the regrouped term list and the pairwise schedule from the alpha table
// (main.c:564-1063): main triples i <= j <= K with alpha != 0, plus the two
// Beta terms -alpha012 t012 g_2 [f0 f1] and +alpha012 t012 g_1 [f0 f2].
struct Schedule {
  std::vector<int> modes;                 // distinct field modes, ascending
  std::vector<int> kmodes;                // distinct outer indices, ascending
  std::vector<NgTerm> terms;              // slots: i,j -> mode slot; grouped by K slot
  std::vector<int> groups;                // K+1 offsets
  std::vector<PairWork> pairs;            // pairwise schedule
  std::vector<NgWeight> weights;
  NgFactored fz;                          // terms factored for the fused z pass
};

// Greedy vertex cover of each group's pairs: repeatedly take the mode slot in the most
// remaining pairs (the lowest slot on ties) and give it all of them.
NgFactored ng_factor(const std::vector<NgTerm>& terms, const std::vector<int>& groups) {
  std::vector<NgRec> recs;
  NgFactored out;
  out.groups.push_back(0);
  for (size_t kk = 0; kk + 1 < groups.size(); ++kk) {
    std::vector<NgTerm> left(terms.begin() + groups[kk], terms.begin() + groups[kk + 1]);
    while (!left.empty()) {
      std::map<int, int> deg;
      for (const NgTerm& t : left) {
        ++deg[t.i];
        if (t.j != t.i) ++deg[t.j];
      }
      int v = left[0].i, best = -1;
      for (auto& [m, d] : deg)
        if (d > best) v = m, best = d;
      const size_t h = recs.size();
      recs.push_back({v, 0, 0.0});
      std::vector<NgTerm> rest;
      for (const NgTerm& t : left) {
        if (t.i == v || t.j == v) {
          const float cf = (float)t.c;
          int bits;
          memcpy(&bits, &cf, 4);
          recs.push_back({t.i == v ? t.j : t.i, bits, t.c});
        } else {
          rest.push_back(t);
        }
      }
      recs[h].aux = (int)(recs.size() - h - 1);
      while (recs.size() % 4) recs.push_back({0, 0, 0.0});
      left.swap(rest);
    }
    out.groups.push_back((int)recs.size());
  }
  out.recs.resize(recs.size() / 4);
  for (size_t q = 0; q < recs.size(); ++q) out.recs[q / 4].q[q % 4] = recs[q];
  return out;
}

Schedule build_schedule(const aar_params& p, const double* alpha) {
  const int Nm = p.N_modes;
  auto A = [&](int a, int b, int c) { return alpha[((size_t)a * Nm + b) * Nm + c]; };
  double t012 = -A(1, 1, 1) / A(0, 1, 2);
  if (A(0, 1, 2) == 0) t012 = 0;
  struct Tri {
    int i, j, K;
    double c;
  };
  std::vector<Tri> tris;
  for (int i = 0; i < Nm; ++i)
    for (int j = i; j < Nm; ++j)
      for (int K = j; K < Nm; ++K)
        if (A(i, j, K) != 0) tris.push_back({i, j, K, A(i, j, K)});
  std::vector<Tri> beta;
  if (t012 != 0) {  // main.c:1043-1052
    beta.push_back({0, 1, 2, -A(0, 1, 2) * t012});
    beta.push_back({0, 2, 1, A(0, 1, 2) * t012});
  }
  Schedule s;
  std::set<int> ms, ks;
  for (auto* v : {&tris, &beta})
    for (auto& t : *v) {
      ms.insert(t.i);
      ms.insert(t.j);
      ks.insert(t.K);
    }
  s.modes.assign(ms.begin(), ms.end());
  s.kmodes.assign(ks.begin(), ks.end());
  std::map<int, int> mslot, kslot;
  for (size_t q = 0; q < s.modes.size(); ++q) mslot[s.modes[q]] = (int)q;
  for (size_t q = 0; q < s.kmodes.size(); ++q) kslot[s.kmodes[q]] = (int)q;
  std::vector<std::vector<NgTerm>> byk(s.kmodes.size());
  for (auto* v : {&tris, &beta})
    for (auto& t : *v) byk[kslot[t.K]].push_back({mslot[t.i], mslot[t.j], t.c});
  s.groups.push_back(0);
  for (auto& v : byk) {
    for (auto& t : v) s.terms.push_back(t);
    s.groups.push_back((int)s.terms.size());
  }
  // pairwise schedule in the reference's loop order: main pairs, then Beta pairs
  for (int i = 0; i < Nm; ++i)
    for (int j = i; j < Nm; ++j) {
      PairWork pw{i, j, (int)s.weights.size(), 0};
      for (int K = j; K < Nm; ++K)
        if (A(i, j, K) != 0) s.weights.push_back({K, A(i, j, K)});
      pw.w_n = (int)s.weights.size() - pw.w_off;
      if (pw.w_n) s.pairs.push_back(pw);
    }
  for (auto& t : beta) {
    PairWork pw{t.i, t.j, (int)s.weights.size(), 1};
    s.weights.push_back({t.K, t.c});
    s.pairs.push_back(pw);
  }
  s.fz = ng_factor(s.terms, s.groups);
  return s;
}

hipEvent_t next_event(aar_job* j) {
  if (j->evnext == j->evpool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    j->evpool.push_back(e);
  }
  return j->evpool[j->evnext++];
}

// Run one launcher of `dispatches` kernel launches; with profiling on, bracket it with events on `st`.
template <class Fn> int timed(aar_job* j, hipStream_t st, int kind, double bytes, Fn&& fn, int dispatches = 1) {
  if (!j->profiling) return fn();
  hipEvent_t a = next_event(j), b = next_event(j);
  if (!a || !b) return AAR_ERR_HIP;
  hipEventRecord(a, st);
  int rc = fn();
  hipEventRecord(b, st);
  j->recs.push_back({kind, bytes, a, b, dispatches});
  return rc;
}

template <typename T> struct Exec {
  using C2 = cx_t<T>;
  static C2* F(aar_job* j, int s) { return (C2*)j->slot(s); }
  static const C2* tw(aar_job* j) {
    if constexpr (sizeof(T) == 4) return j->ctx->tw32;
    else return j->ctx->tw64;
  }

  static int run(aar_job* j, hipStream_t st) {
    using P = Pipeline<T>;
    const Geom& g = j->g;
    const aar_params& p = j->p;
    const aar_cosmo& c = j->c;
    int rc;
    const double S = (double)g.fs * sizeof(C2);              // one spectrum / padded field
    const double R = (double)g.N * g.N * g.N * sizeof(T);     // one real field
    const double phase_shift = p.PhaseFlip == 1 ? AAR_PI : 0.0;
    // x / y pass pairs, in the tuned column groups; one timed unit of `dispatches` launches
    const int xyg = aar_xy_group(sizeof(T) == 8, g.N);
    const int xsg = aar_xy_split_group(sizeof(T) == 8, g.N);
    auto xy = [&](C2* f, int nf, int dir) {
      return timed(j, st, AAR_K_XY, 4 * nf * S, [&] { return P::xy_pair(f, nf, g, tw(j), dir, st, xyg); },
                   aar_xy_pair_dispatches(nf, g.H, xyg));
    };
    auto c2r = [&](C2* f, int nf) {
      int r;
      if ((r = xy(f, nf, +1))) return r;
      return timed(j, st, AAR_K_Z, nf * (S + R), [&] { return P::z_pass(f, nf, g, tw(j), +1, st); });
    };
    auto r2c = [&](C2* f, int nf) {
      int r;
      if ((r = timed(j, st, AAR_K_Z, nf * (S + R), [&] { return P::z_pass(f, nf, g, tw(j), -1, st); })))
        return r;
      return xy(f, nf, -1);
    };
    if ((rc = timed(j, st, AAR_K_GAUSS, S + 4.0 * g.N * g.N, [&] {
           return P::gauss(F(j, j->s_phig), g, j->ks, j->seeds, c.Anorm, p.PrimordialIndex, c.fac, c.Beta,
                           phase_shift, p.FixedAmplitude, st);
         })))
      return rc;
    C2* reg = F(j, j->s_reg);
    const double n3 = (double)((unsigned)g.N * (unsigned)g.N * (unsigned)g.N);
    if (j->sched == AAR_SCHED_BATCHED) {
      if (j->M > 0) {
        const bool fused = j->M <= 16 && g.N >= 64 && g.pow2;
        const bool modes_x = fused && (j->ctx->modes_x == 2 || (j->ctx->modes_x == 1 && j->short_filters));
        if (modes_x) {
          // the mode fields filtered from phi_G in the x pass's loads, then their y passes
          if ((rc = timed(j, st, AAR_K_MODES_XY, (1 + 3.0 * j->M) * S,
                          [&] { return P::modes_xy(F(j, j->s_phig), reg, j->M, j->modes, g, j->ks, tw(j), st, xyg); },
                          aar_xy_pair_dispatches(1, g.H, xyg))))
            return rc;
        } else if ((rc = timed(j, st, AAR_K_MODES, (1 + j->M) * S,
                               [&] { return P::modes_spec(F(j, j->s_phig), reg, j->M, j->modes, g, j->ks, st); }))) {
          return rc;
        }
        if (fused) {
          // x, y inverse passes of the M mode fields, then the fused z pass
          // (C2R, products, R2C), then y, x forward passes of the K products
          if (!modes_x && (rc = xy(reg, j->M, +1))) return rc;
          if ((rc = timed(j, st, AAR_K_PRODUCTS, (j->M + j->K) * S, [&] {
                 return j->jit.fn ? jit_zng_launch(j->jit, reg, j->M, j->K, g, tw(j), st)
                                  : P::z_ng(reg, j->M, j->K, j->fgroups, j->frecs, g, tw(j), st);
               })))
            return rc;
          if ((rc = xy(reg, j->K, -1))) return rc;
        } else {
          if ((rc = c2r(reg, j->M))) return rc;
          if ((rc = timed(j, st, AAR_K_PRODUCTS, (j->M + j->K) * R,
                          [&] { return P::products(reg, j->M, reg, j->K, j->terms, j->groups, g, st); })))
            return rc;
          if ((rc = r2c(reg, j->K))) return rc;
        }
      }
      // combine + the nine 2LPT spectra in one elementwise pass (with the fused z passes: only
      // the five x-pass inputs of the nine, the ky and kz factors applied by the x pass)
      const bool split = j->fused_lpt;
      if ((rc = timed(j, st, AAR_K_COMBINE, (j->K + (split ? 7 : 11)) * S, [&] {
             return P::ng_lpt(F(j, j->s_phig), reg, j->K, j->kmodes, F(j, j->s_phi), split ? reg + 9 * g.fs : reg,
                              p.Fnl, g, j->ks, j->tf, c.Dplus, c.Beta, st, split);
           })))
        return rc;
      C2* disp = F(j, j->s_disp);
      C2* disp2 = F(j, j->s_disp2);
      // x then y inverse passes of fields made by XSplit xs: nin x-pass inputs -> nout spectra
      auto xys = [&](C2* f, const XSplit& xs) {
        return timed(j, st, AAR_K_XY_SPLIT, (3.0 * xs.nin + xs.nout) * S,
                     [&] { return P::xy_split(f, xs, g, tw(j), st, xsg); }, aar_xy_pair_dispatches(xs.nin, g.H, xsg));
      };
      if (split) {
        // x (five inputs, nine outputs), y inverse passes of the nine spectra; z C2R of the
        // three ZA fields; the six gradient fields go through the fused z pass, whose one
        // product is the 2LPT source (never stored in real space), then y, x forward passes
        // of the source; its two x-pass inputs of the three 2LPT displacements
        if ((rc = xys(reg, lpt_xsplit(j->ks)))) return rc;
        if ((rc = timed(j, st, AAR_K_Z, 3 * (S + R), [&] { return P::z_pass(disp, 3, g, tw(j), +1, st); })))
          return rc;
        if ((rc = timed(j, st, AAR_K_PRODUCTS, 7 * S,
                        [&] { return P::z_source(reg, g, tw(j), st); })))
          return rc;
        if ((rc = xy(reg, 1, -1))) return rc;
        if ((rc = timed(j, st, AAR_K_LPT, 3 * S,
                        [&] { return P::disp2_spec(reg, disp2 + 3 * g.fs, g, j->ks, st, true); })))
          return rc;
        if ((rc = xys(disp2, disp2_xsplit(j->ks)))) return rc;
        if ((rc = timed(j, st, AAR_K_Z, 3 * (S + R), [&] { return P::z_pass(disp2, 3, g, tw(j), +1, st); })))
          return rc;
      } else {
        if ((rc = c2r(reg, 9))) return rc;
        if ((rc = timed(j, st, AAR_K_PRODUCTS, 7 * R, [&] { return P::source(reg, g, st); }))) return rc;
        if ((rc = r2c(reg + 3 * g.fs, 1))) return rc;
        if ((rc = timed(j, st, AAR_K_LPT, 4 * S, [&] { return P::disp2_spec(reg + 3 * g.fs, disp2, g, j->ks, st); })))
          return rc;
        if ((rc = c2r(disp2, 3))) return rc;
      }
      return timed(j, st, AAR_K_CIC, 6 * R + 24.0 * j->npart, [&] {
        return P::cic(disp, disp2, g, j->glass, j->nglass, j->glass_box, j->p.GlassTileFac, p.Box,
                      c.vel_prefac, c.vel_prefac2, j->pos, j->vel, j->npart, st);
      });
    } else {
      if ((rc = aar_zero_fill((double*)j->acc, 2 * g.fs, st))) return rc;
      C2* fi = reg;
      C2* fj = F(j, j->s_reg + 1);
      int cur_i = -1;
      for (const PairWork& pw : j->pairs) {
        if (pw.i != cur_i) {  // main.c:571-652: build and C2R the field of N_i
          if ((rc = timed(j, st, AAR_K_MODES, 2 * S,
                          [&] { return P::modes_spec(F(j, j->s_phig), fi, 1, j->modes + pw.i, g, j->ks, st); })))
            return rc;
          if ((rc = c2r(fi, 1))) return rc;
          cur_i = pw.i;
        }
        // main.c:666-781: field of N_j, C2R, product, R2C
        if ((rc = timed(j, st, AAR_K_MODES, 2 * S,
                        [&] { return P::modes_spec(F(j, j->s_phig), fj, 1, j->modes + pw.j, g, j->ks, st); })))
          return rc;
        if ((rc = c2r(fj, 1))) return rc;
        if ((rc = timed(j, st, AAR_K_PRODUCTS, 3 * R,
                        [&] { return P::products(fi, 2, fj, 1, j->terms, j->groups, g, st); })))
          return rc;
        if ((rc = r2c(fj, 1))) return rc;
        // main.c:785-870: weighted accumulation
        if ((rc = timed(j, st, AAR_K_ACCUM, S + 2.0 * g.fs * 16, [&] {
               return P::ng_accumulate(j->acc, fj, j->weights + pw.w_off, pw.w_n, g, j->ks, st);
             })))
          return rc;
      }
      if ((rc = timed(j, st, AAR_K_COMBINE, 2 * S + g.fs * 16.0, [&] {
             return P::ng_combine_acc(F(j, j->s_phig), j->acc, F(j, j->s_phi), p.Fnl, n3, g, st);
           })))
        return rc;
    }
    // 2LPT (main.c:1158-1381)
    if ((rc = timed(j, st, AAR_K_LPT, 10 * S,
                    [&] { return P::lpt_spec(F(j, j->s_phi), reg, g, j->ks, j->tf, c.Dplus, c.Beta, st); })))
      return rc;
    if ((rc = c2r(reg, 9))) return rc;
    if ((rc = timed(j, st, AAR_K_PRODUCTS, 7 * R, [&] { return P::source(reg, g, st); }))) return rc;
    if ((rc = r2c(reg + 3 * g.fs, 1))) return rc;
    if ((rc = timed(j, st, AAR_K_LPT, 4 * S, [&] { return P::disp2_spec(reg + 3 * g.fs, reg, g, j->ks, st); })))
      return rc;
    if ((rc = c2r(reg, 3))) return rc;
    // CIC readout and particle update (main.c:1440-1516)
    return timed(j, st, AAR_K_CIC, 6 * R + 24.0 * j->npart, [&] {
      return P::cic(reg + 6 * g.fs, reg, g, j->glass, j->nglass, j->glass_box, j->p.GlassTileFac, p.Box,
                    c.vel_prefac, c.vel_prefac2, j->pos, j->vel, j->npart, st);
    });
  }
};

// Reference layouts (row x*N + y, kz contiguous, N/2+1 bins) <-> the x-blocked
// device layout (aar_common.h).
template <typename T> void to_internal_spec(const double* ref, T* out, int N) {
  const int H = N / 2;
  const long long nq = (long long)N * N * H;
  for (int x = 0; x < N; ++x)
    for (int y = 0; y < N; ++y) {
      const long long q = (long long)x * N + y, r = row_of(N, x, y);
      for (int k = 0; k < H; ++k) {
        out[2 * (r * H + k)] = (T)ref[2 * (q * (H + 1) + k)];
        out[2 * (r * H + k) + 1] = (T)ref[2 * (q * (H + 1) + k) + 1];
      }
      out[2 * (nq + r)] = (T)ref[2 * (q * (H + 1) + H)];
      out[2 * (nq + r) + 1] = (T)ref[2 * (q * (H + 1) + H) + 1];
    }
}

template <typename T> void from_internal_spec(const T* in, double* ref, int N) {
  const int H = N / 2;
  const long long nq = (long long)N * N * H;
  for (int x = 0; x < N; ++x)
    for (int y = 0; y < N; ++y) {
      const long long q = (long long)x * N + y, r = row_of(N, x, y);
      for (int k = 0; k < H; ++k) {
        ref[2 * (q * (H + 1) + k)] = in[2 * (r * H + k)];
        ref[2 * (q * (H + 1) + k) + 1] = in[2 * (r * H + k) + 1];
      }
      ref[2 * (q * (H + 1) + H)] = in[2 * (nq + r)];
      ref[2 * (q * (H + 1) + H) + 1] = in[2 * (nq + r) + 1];
    }
}

template <typename T> void to_internal_real(const double* ref, T* out, int N) {
  for (int x = 0; x < N; ++x)
    for (int y = 0; y < N; ++y) {
      const long long q = ((long long)x * N + y) * N, r = row_of(N, x, y) * N;
      for (int z = 0; z < N; ++z) out[r + z] = (T)ref[q + z];
    }
}

template <typename T> void from_internal_real(const T* in, double* ref, int N) {
  for (int x = 0; x < N; ++x)
    for (int y = 0; y < N; ++y) {
      const long long q = ((long long)x * N + y) * N, r = row_of(N, x, y) * N;
      for (int z = 0; z < N; ++z) ref[q + z] = in[r + z];
    }
}

int device_alloc(void** p, size_t bytes) {
  if (hipMalloc(p, bytes ? bytes : 1) != hipSuccess) {
    *p = nullptr;
    return AAR_ERR_ALLOC;
  }
  return AAR_OK;
}

template <typename V> int upload(V** dst, const std::vector<V>& src) {
  int rc = device_alloc((void**)dst, sizeof(V) * src.size());
  if (rc) return rc;
  if (!src.empty() && hipMemcpy(*dst, src.data(), sizeof(V) * src.size(), hipMemcpyHostToDevice) != hipSuccess)
    return AAR_ERR_HIP;
  return AAR_OK;
}

// Device memory of a job: its own hipMalloc'd buffer, or the next 256-byte
// aligned piece of the workspace it was prepared in.
int job_alloc(aar_job* j, void** p, size_t bytes) {
  if (!j->ws) return device_alloc(p, bytes);
  const size_t off = (j->ws_used + 255) & ~(size_t)255, n = bytes ? bytes : 1;
  if (off + n > j->ws_size) {
    *p = nullptr;
    return AAR_ERR_HBM_BUDGET;
  }
  *p = j->ws + off;
  j->ws_used = off + n;
  return AAR_OK;
}

template <typename V> int job_upload(aar_job* j, V** dst, const std::vector<V>& src) {
  const size_t n = sizeof(V) * src.size();
  int rc = job_alloc(j, (void**)dst, n);
  if (rc || !n) return rc;
  if (!j->up) return hipMemcpy(*dst, src.data(), n, hipMemcpyHostToDevice) == hipSuccess ? AAR_OK : AAR_ERR_HIP;
  j->staged.emplace_back((const char*)src.data(), (const char*)src.data() + n);
  return hipMemcpyAsync(*dst, j->staged.back().data(), n, hipMemcpyHostToDevice, j->up) == hipSuccess
             ? AAR_OK
             : AAR_ERR_HIP;
}

}  // namespace

// ------------------------------------------------------------------ scratch
// Device memory of the allocation-free entry point (aar_job_run_noalloc): every buffer of a job
// besides its inputs and outputs (tables, seed table, field slots) is carved from this array.
// Each GPU the library runs on holds its own copy until the process ends.  Host code never
// passes the array's name to a kernel or a copy: it uses the address hipGetSymbolAddress
// returned for the device at set-up (g_scratch_ptr).
static constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) char scratch[kScratchBytes];
static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

// With `device` current: look up this device's copy of the scratch array once.
static int scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return AAR_ERR_BAD_ARG;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return AAR_OK;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    fprintf(stderr, "hipGetSymbolAddress(scratch) failed on device %d: %s\n", device, hipGetErrorString(e));
    return AAR_ERR_HIP;
  }
  g_scratch_ptr[device] = ptr;
  return AAR_OK;
}

// ------------------------------------------------------------------ context
extern "C" int aar_ctx_create(int device, aar_ctx** out) {
  if (!out) return AAR_ERR_BAD_ARG;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0) return AAR_ERR_NO_DEVICE;
  AAR_HIP(hipSetDevice(device));
  const int src = scratch_setup(device);
  if (src) return src;
  aar_ctx* c = new aar_ctx();
  c->device = device;
  AAR_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  std::vector<float2> t32(kTwTotal);
  std::vector<double2> t64(kTwTotal);
  make_twiddles<float>(t32.data());
  make_twiddles<double>(t64.data());
  int rc = upload(&c->tw32, t32);
  if (!rc) rc = upload(&c->tw64, t64);
  if (rc) {
    delete c;
    return rc;
  }
  *out = c;
  return AAR_OK;
}

int aar_ctx_device(const aar_ctx* c) { return c->device; }
hipStream_t aar_ctx_stream(const aar_ctx* c) { return c->stream; }
int aar_device_alloc(void** p, size_t bytes) { return device_alloc(p, bytes); }

extern "C" int aar_ctx_sync(aar_ctx* c) {
  if (!c) return AAR_ERR_BAD_ARG;
  AAR_HIP(hipSetDevice(c->device));
  AAR_HIP(hipDeviceSynchronize());
  return AAR_OK;
}

extern "C" int aar_ctx_destroy(aar_ctx* c) {
  if (!c) return AAR_OK;
  hipSetDevice(c->device);
  c->plans.clear();  // jobs hold their plans until they are destroyed
  hipFree(c->tw32);
  hipFree(c->tw64);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return AAR_OK;
}

// ------------------------------------------------------------------ autotuner
// Times every tile variant of the x and y passes on nf fields (alternating
// inverse and forward launches) and keeps the fastest per axis in the
// variant table.  The variants share the radix plan, so the choice changes
// speed only, never results.
static std::mutex g_tune_mu;
static bool g_tuned[2][11] = {};

template <typename T>
static int autotune_t(aar_ctx* ctx, int N, int nf, void* scratch, int nslots, double* sec, int* chosen,
                      double* plan_sec = nullptr, double* split_sec = nullptr, double* pair_sec = nullptr) {
  using C2 = cx_t<T>;
  const Geom g = make_geom(N);
  C2* d = (C2*)scratch;
  // fields of d: nf for the pass timings, at least fourteen for the multi-output pairs (scratch,
  // the job's own pool, holds nslots)
  const int nall = scratch ? nslots : std::max(nf, 14);
  if (!d && device_alloc((void**)&d, (size_t)nall * g.fs * sizeof(C2))) return AAR_ERR_ALLOC;
  int rc = aar_noise_fill<T>((T*)d, (long long)nall * g.fs * 2, ctx->stream);
  const C2* tw = sizeof(T) == 4 ? (const C2*)ctx->tw32 : (const C2*)ctx->tw64;
  hipEvent_t a = nullptr, b = nullptr;
  hipEventCreate(&a);
  hipEventCreate(&b);
  const int iters = 4;
  bool built_y[kXYVariantCount] = {};  // tile shapes built for the y pass at this size
  for (int axis = 0; axis < 2 && !rc; ++axis) {
    double best = 1e30;
    int bv = 0;
    for (int v = 0; v < kXYVariantCount && !rc; ++v) {
      double t = 1e30;
      int r = Pipeline<T>::xy_pass(d, nf, g, tw, axis, +1, ctx->stream, v);
      if (axis == 1) built_y[v] = r == AAR_OK;
      if (r == AAR_OK) {
        hipEventRecord(a, ctx->stream);
        for (int q = 0; q < iters && !r; ++q) r = Pipeline<T>::xy_pass(d, nf, g, tw, axis, q % 2 ? -1 : +1, ctx->stream, v);
        hipEventRecord(b, ctx->stream);
        if (hipEventSynchronize(b) != hipSuccess) r = AAR_ERR_HIP;
        float ms = 0;
        if (!r && hipEventElapsedTime(&ms, a, b) == hipSuccess) t = ms * 1e-3 / iters;
      }
      if (r && r != AAR_ERR_UNSUPPORTED) rc = r;
      if (sec) sec[axis * kXYVariantCount + v] = t < 1e30 ? t : -1.0;
      if (t < best) {
        best = t;
        bv = v;
      }
    }
    if (!rc) aar_set_xy_variant(sizeof(T) == 8, N, axis, bv);
    if (chosen) chosen[axis] = bv;
  }
  // then, jointly, the radix plan of the x / y passes (n = 512, 1024: three stages in different
  // radix orders with the same tile shapes; plans differ in rounding only) and the column groups
  // of the x / y pass pairs (whole passes over the nf fields, or field by field over all kz
  // columns, halves or quarters), each pair timed both ways
  const int nplan = (N == 512 || N == 1024) ? kXYPlanCount : 1;
  double best = 1e30;
  int bg = 0, bp = 0;
  // seconds per call of fn(q) (q = 0, 1, ...): one call untimed, then the least mean of three
  // runs of `iters` calls (the choices below are a few percent apart)
  auto time_calls = [&](auto&& fn, double& t) {
    t = 1e30;
    int r = fn(0);
    for (int rep = 0; rep < 3 && !r; ++rep) {
      hipEventRecord(a, ctx->stream);
      for (int q = 0; q < iters && !r; ++q) r = fn(q);
      hipEventRecord(b, ctx->stream);
      if (hipEventSynchronize(b) != hipSuccess) r = AAR_ERR_HIP;
      float ms = 0;
      if (!r && hipEventElapsedTime(&ms, a, b) == hipSuccess) t = std::min(t, ms * 1e-3 / iters);
    }
    return r;
  };
  for (int pl = 0; pl < nplan && !rc; ++pl) {
    aar_set_xy_plan(sizeof(T) == 8, N, pl);
    double bestp = 1e30;
    int gi = 0;
    for (int gc : {0, g.H, g.H / 2, g.H / 4}) {
      if (rc || (gc > 0 && gc < 16)) break;
      double t;
      const int r = time_calls(
          [&](int q) { return Pipeline<T>::xy_pair(d, nf, g, tw, q % 2 ? -1 : +1, ctx->stream, gc); }, t);
      if (r) rc = r;
      if (pair_sec) pair_sec[pl * 4 + gi] = t < 1e30 ? t : -1.0;
      ++gi;
      bestp = std::min(bestp, t);
      if (t < best) {
        best = t;
        bg = gc;
        bp = pl;
      }
    }
    if (plan_sec) plan_sec[pl] = bestp < 1e30 ? bestp : -1.0;
  }
  aar_set_xy_plan(sizeof(T) == 8, N, rc ? 0 : bp);
  if (!rc) aar_set_xy_group(sizeof(T) == 8, N, bg);
  // the column groups of the multi-output pairs of the 2LPT spectra (five inputs, nine outputs:
  // a group's outputs are larger than its input, so the best group can differ), on fourteen
  // fields with unit factor tables
  double* ones = nullptr;
  if (!rc && nall >= 14 && !device_alloc((void**)&ones, sizeof(double) * N)) {
    const std::vector<double> one(N, 1.0);
    XSplit xs = lpt_xsplit(KSpace{});
    xs.kv = xs.icic = ones;
    double bests = 1e30;
    int bs = bg;
    if (hipMemcpy(ones, one.data(), sizeof(double) * N, hipMemcpyHostToDevice) != hipSuccess) rc = AAR_ERR_HIP;
    aar_set_xy_variant(sizeof(T) == 8, N, 2, -1);  // the groups timed with the y pass's tile shape
    for (int gc : {0, g.H, g.H / 2, g.H / 4, g.H / 8}) {
      if (rc || (gc > 0 && gc < 16)) break;
      double t;
      const int r = time_calls([&](int) { return Pipeline<T>::xy_split(d, xs, g, tw, ctx->stream, gc); }, t);
      if (r) rc = r;
      if (split_sec) split_sec[gc ? 1 + __builtin_ctz(g.H / gc) : 0] = t < 1e30 ? t : -1.0;
      if (t < bests) {
        bests = t;
        bs = gc;
      }
    }
    if (!rc) aar_set_xy_split_group(sizeof(T) == 8, N, bs);
    // then the tile shape of the pairs' y pass (its loads apply the output factors, one output
    // per workgroup, so the shape that suits the plain y pass need not suit it), at that group
    double bestv = 1e30;
    int bv = -1;
    for (int v = 0; v < kXYVariantCount && !rc; ++v) {
      if (!built_y[v]) continue;  // not built here: it would run as shape 0
      aar_set_xy_variant(sizeof(T) == 8, N, 2, v);
      double t = 1e30;
      const int r = time_calls([&](int) { return Pipeline<T>::xy_split(d, xs, g, tw, ctx->stream, bs); }, t);
      if (r) rc = r;
      if (split_sec) split_sec[5 + v] = t < 1e30 ? t : -1.0;
      if (t < bestv) {
        bestv = t;
        bv = v;
      }
    }
    aar_set_xy_variant(sizeof(T) == 8, N, 2, rc ? -1 : bv);
    hipFree(ones);
  }
  hipEventDestroy(a);
  hipEventDestroy(b);
  if (!scratch) hipFree(d);
  else if (!rc)  // a job's pool: leave it as a fresh allocation is left to the job (zeroed)
    rc = aar_zero_fill((double*)d, (long long)nall * g.fs * (long long)(sizeof(C2) / 8), ctx->stream);
  if (!rc) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned[sizeof(T) == 8][ilog2i(N)] = true;
  }
  return rc;
}

extern "C" int aar_xy_group_cols(int n, int precision) {
  if (!pow2(n) || (precision != AAR_FP32 && precision != AAR_FP64)) return AAR_ERR_BAD_ARG;
  return aar_xy_group(precision == AAR_FP64, n);
}

extern "C" int aar_set_xy_group_cols(int n, int precision, int cols) {
  if (!pow2(n) || (precision != AAR_FP32 && precision != AAR_FP64)) return AAR_ERR_BAD_ARG;
  if (cols < 0 || cols > n / 2 || (cols && cols % 16)) return AAR_ERR_BAD_ARG;
  aar_set_xy_group(precision == AAR_FP64, n, cols);
  return AAR_OK;
}

extern "C" int aar_autotune(aar_ctx* ctx, int n, int precision, int nf, double* seconds, int* variant) {
  if (!ctx || !pow2(n) || nf < 1 || (precision != AAR_FP32 && precision != AAR_FP64)) return AAR_ERR_BAD_ARG;
  AAR_HIP(hipSetDevice(ctx->device));
  return precision == AAR_FP32 ? autotune_t<float>(ctx, n, nf, nullptr, 0, seconds, variant)
                               : autotune_t<double>(ctx, n, nf, nullptr, 0, seconds, variant);
}

extern "C" int aar_autotune_plans(aar_ctx* ctx, int n, int precision, int nf, double* plan_seconds, int* plan) {
  if (!ctx || !pow2(n) || nf < 1 || (precision != AAR_FP32 && precision != AAR_FP64)) return AAR_ERR_BAD_ARG;
  AAR_HIP(hipSetDevice(ctx->device));
  if (plan_seconds)
    for (int q = 0; q < kXYPlanCount; ++q) plan_seconds[q] = -1.0;
  const int rc = precision == AAR_FP32 ? autotune_t<float>(ctx, n, nf, nullptr, 0, nullptr, nullptr, plan_seconds)
                                       : autotune_t<double>(ctx, n, nf, nullptr, 0, nullptr, nullptr, plan_seconds);
  if (plan) *plan = aar_xy_plan(precision == AAR_FP64, n);
  return rc;
}

extern "C" int aar_autotune_report(aar_ctx* ctx, int n, int precision, int nf, double* tile_seconds,
                                   double* pair_seconds, double* split_seconds) {
  if (!ctx || !pow2(n) || nf < 1 || (precision != AAR_FP32 && precision != AAR_FP64)) return AAR_ERR_BAD_ARG;
  AAR_HIP(hipSetDevice(ctx->device));
  if (pair_seconds)
    for (int q = 0; q < 4 * kXYPlanCount; ++q) pair_seconds[q] = -1.0;
  if (split_seconds)
    for (int q = 0; q < 9; ++q) split_seconds[q] = -1.0;
  return precision == AAR_FP32
             ? autotune_t<float>(ctx, n, nf, nullptr, 0, tile_seconds, nullptr, nullptr, split_seconds, pair_seconds)
             : autotune_t<double>(ctx, n, nf, nullptr, 0, tile_seconds, nullptr, nullptr, split_seconds, pair_seconds);
}

extern "C" int aar_xy_split_group_cols(int n, int precision) {
  if (!pow2(n) || (precision != AAR_FP32 && precision != AAR_FP64)) return AAR_ERR_BAD_ARG;
  return aar_xy_split_group(precision == AAR_FP64, n);
}

extern "C" int aar_set_xy_split_group_cols(int n, int precision, int cols) {
  if (!pow2(n) || (precision != AAR_FP32 && precision != AAR_FP64)) return AAR_ERR_BAD_ARG;
  if (cols < -1 || cols > n / 2 || (cols > 0 && cols % 16)) return AAR_ERR_BAD_ARG;
  aar_set_xy_split_group(precision == AAR_FP64, n, cols);
  return AAR_OK;
}

extern "C" int aar_xy_radix_plan(int n, int precision) {
  if (!pow2(n) || (precision != AAR_FP32 && precision != AAR_FP64)) return AAR_ERR_BAD_ARG;
  return aar_xy_plan(precision == AAR_FP64, n);
}

extern "C" int aar_set_xy_radix_plan(int n, int precision, int plan) {
  if (!pow2(n) || (precision != AAR_FP32 && precision != AAR_FP64)) return AAR_ERR_BAD_ARG;
  if (plan < 0 || plan >= kXYPlanCount || (plan && n != 512 && n != 1024)) return AAR_ERR_BAD_ARG;
  aar_set_xy_plan(precision == AAR_FP64, n, plan);
  return AAR_OK;
}

extern "C" int aar_set_xy_tile_variant(int n, int precision, int axis, int variant) {
  if (!pow2(n) || axis < 0 || axis > 2 || variant < (axis == 2 ? -1 : 0) || variant >= kXYVariantCount)
    return AAR_ERR_BAD_ARG;
  if (precision != AAR_FP32 && precision != AAR_FP64) return AAR_ERR_BAD_ARG;
  aar_set_xy_variant(precision == AAR_FP64, n, axis, variant);
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tuned[precision == AAR_FP64][ilog2i(n)] = true;  // a pinned choice is not re-tuned
  return AAR_OK;
}

extern "C" int aar_xy_tile_variant(int n, int precision, int axis) {
  if (!pow2(n) || axis < 0 || axis > 2) return -1;
  return aar_xy_variant(precision == AAR_FP64, n, axis);
}

// ------------------------------------------------------------------ jobs
// aar_cosmology through the context's cache.  The key holds every parameter
// the cosmology code reads (and a few more) plus the transfer table rows.
static int ctx_cosmology(aar_ctx* ctx, const aar_params* p, const aar_tables* t, aar_cosmo* c) {
  std::string key;
  auto put = [&](const void* v, size_t n) { key.append((const char*)v, n); };
  const double d[] = {p->Box, p->Omega, p->OmegaLambda, p->OmegaBaryon, p->OmegaDM_2ndSpecies, p->HubbleParam,
                      p->Sigma8, p->PrimordialIndex, p->Redshift, p->ShapeGamma, p->UnitLength_in_cm,
                      p->UnitMass_in_g, p->UnitVelocity_in_cm_per_s, p->InputSpectrum_UnitLength_in_cm,
                      p->WDM_PartMass_in_kev, p->A_res, p->w_res, p->phi_res};
  const int i[] = {p->WhichSpectrum, p->WhichTransfer, p->WDM_On, p->WDM_Vtherm_On, p->Variant};
  put(d, sizeof d);
  put(i, sizeof i);
  if (p->WhichTransfer == 2 && t->ntransfer > 0 && t->transfer_k && t->transfer_t) {
    put(t->transfer_k, sizeof(double) * t->ntransfer);
    put(t->transfer_t, sizeof(double) * t->ntransfer);
  }
  {
    std::lock_guard<std::mutex> g(ctx->mu);
    auto it = ctx->cosmo_cache.find(key);
    if (it != ctx->cosmo_cache.end()) {
      *c = it->second;
      return AAR_OK;
    }
  }
  const int rc = aar_cosmology(p, t, c);
  if (rc) return rc;
  std::lock_guard<std::mutex> g(ctx->mu);
  if (ctx->cosmo_cache.size() > 256) ctx->cosmo_cache.clear();
  ctx->cosmo_cache[key] = *c;
  return AAR_OK;
}

static int validate(const aar_params* p, const aar_tables* t) {
  if (!p || !t) return AAR_ERR_BAD_ARG;
  if (!mesh_ok(p->Nmesh)) return AAR_ERR_UNSUPPORTED;
  if (p->Nmax < p->Nsample) return AAR_ERR_NMAX;
  if (p->N_modes < 3 || !t->alpha || !t->x0) return AAR_ERR_BAD_ARG;
  if (!t->glass || t->nglass <= 0 || p->GlassTileFac <= 0) return AAR_ERR_BAD_ARG;
  // The CIC gather indexes the fields with the cell of each particle's unperturbed position
  // (main.c:1443-1461 clamps only i == N, j == N, k == N and i < 0): a glass coordinate outside
  // [0, BoxSize] or not finite, or a box that is not positive, would index outside the grid.
  if (!(p->Box > 0) || !std::isfinite(p->Box) || !(t->glass_box > 0) || !std::isfinite(t->glass_box))
    return AAR_ERR_BAD_ARG;
  const float gb = (float)t->glass_box;
  for (size_t q = 0; q < 3 * (size_t)t->nglass; ++q) {
    const float v = t->glass[q];
    if (!(v >= 0.0f && v <= gb)) return AAR_ERR_BAD_ARG;  // NaN fails both
  }
  return AAR_OK;
}

// Field slots of a job: phi_G, Phi and a region of the mode fields, the products and the 2LPT
// spectra; with the fused z passes (batched, N >= 64, powers of two) the nine 2LPT spectra and
// the five x-pass inputs they are made from (lpt_xsplit).
static int slots_needed(int M, int K, int sched, int N) {
  if (sched == AAR_SCHED_PLAIN) return 2 + 9;
  const bool fused = N >= 64 && is_pow2(N);
  return 2 + std::max(std::max(M, K), fused ? 14 : 9);
}

// The host half of a plan: schedule, k-space and transfer tables, glass, and the image of the
// device block.  Needs no device (aar_job_bytes sizes a job with it).
static int plan_layout(const aar_params* p, const aar_tables* t, int schedule, Plan& pl) {
  const int N = p->Nmesh, Nm = p->N_modes;
  Schedule s = build_schedule(*p, t->alpha);
  pl.modes_h = s.modes;
  pl.kmodes_h = s.kmodes;
  pl.terms_h = s.terms;
  pl.groups_h = s.groups;
  pl.M = (int)s.modes.size();
  pl.K = (int)s.kmodes.size();
  pl.pairs = s.pairs;
  pl.fgroups_h = s.fz.groups;
  pl.frecs_h = s.fz.recs;
  if (pl.M > 32) return AAR_ERR_UNSUPPORTED;
  std::vector<int> modes_dev = s.modes, groups = s.groups;
  std::vector<NgTerm> terms = s.terms;
  if (schedule == AAR_SCHED_PLAIN) {
    // pairwise: fields are addressed by mode number, one product term f0*f1
    modes_dev.resize(Nm);
    for (int q = 0; q < Nm; ++q) modes_dev[q] = q;
    terms = {NgTerm{0, 1, 1.0}};
    groups = {0, 1};
  }
  // k-space tables (main.c:412-452, 1323-1343)
  std::vector<double> kv(N), sinc(N), icic(N), x0(Nm);
  std::vector<unsigned char> outc(N);
  for (int i = 0; i < N; ++i) {
    kv[i] = (i < N / 2) ? i * 2 * AAR_PI / p->Box : -(N - i) * 2 * AAR_PI / p->Box;
    outc[i] = fabs(kv[i]) * p->Box / (2 * AAR_PI) > p->Nsample / 2;
    double f = 1;
    if (kv[i] != 0) {
      f = (kv[i] * p->Box / 2) / N;
      f = sin(f) / f;
    }
    sinc[i] = f;
    icic[i] = (1 / f) * (1 / f);
  }
  for (int q = 0; q < Nm; ++q) x0[q] = t->x0[2 * q];
  std::vector<double> logk, logT;
  int rc;
  if (p->WhichTransfer == 2 && (rc = aar_host_transfer_tables(p, t, logk, logT))) return rc;
  pl.nt = (int)logk.size();
  // pack every table at a 256-byte aligned offset of one block
  struct Part {
    const void* src;
    size_t n;
    size_t* at;
  };
  auto bytes_of = [](const auto& v) { return sizeof(v[0]) * v.size(); };
  const Part parts[] = {
      {kv.data(), bytes_of(kv), &pl.o_kv},
      {sinc.data(), bytes_of(sinc), &pl.o_sinc},
      {icic.data(), bytes_of(icic), &pl.o_icic},
      {x0.data(), bytes_of(x0), &pl.o_x0},
      {outc.data(), bytes_of(outc), &pl.o_outc},
      {modes_dev.data(), bytes_of(modes_dev), &pl.o_modes},
      {s.kmodes.data(), bytes_of(s.kmodes), &pl.o_kmodes},
      {groups.data(), bytes_of(groups), &pl.o_groups},
      {terms.data(), bytes_of(terms), &pl.o_terms},
      {s.weights.data(), bytes_of(s.weights), &pl.o_weights},
      {s.fz.groups.data(), bytes_of(s.fz.groups), &pl.o_fgroups},
      {s.fz.recs.data(), bytes_of(s.fz.recs), &pl.o_frecs},
      {logk.data(), bytes_of(logk), &pl.o_logk},
      {logT.data(), bytes_of(logT), &pl.o_logT},
      {t->glass, sizeof(float) * 3 * (size_t)t->nglass, &pl.o_glass},
  };
  size_t off = 0;
  for (const Part& q : parts) {
    off = (off + 255) & ~(size_t)255;
    *q.at = off;
    off += q.n ? q.n : 1;
  }
  pl.bytes = (off + 255) & ~(size_t)255;
  pl.image.assign(pl.bytes, 0);
  for (const Part& q : parts)
    if (q.n) memcpy(pl.image.data() + *q.at, q.src, q.n);
  pl.nglass = t->nglass;
  pl.glass_box = t->glass_box;
  return AAR_OK;
}

// Bytes of a job's own device buffers (field slots, pairwise accumulator, seed table,
// positions, velocities), each rounded up to 256 bytes.
static size_t job_own_bytes(const aar_params* p, const aar_tables* t, int precision, int schedule, int M, int K,
                            bool particles = true) {
  const Geom g = make_geom(p->Nmesh);
  const size_t N = p->Nmesh, cx = precision == AAR_FP32 ? 8 : 16;
  const size_t npart = (size_t)p->GlassTileFac * p->GlassTileFac * p->GlassTileFac * t->nglass;
  size_t b = 0;
  auto add = [&](size_t n) { b = ((b + 255) & ~(size_t)255) + (n ? n : 1); };
  add(slots_needed(M, K, schedule, N) * (size_t)g.fs * cx);
  if (schedule == AAR_SCHED_PLAIN) add((size_t)g.fs * 16);
  add(4 * N * N);
  if (particles) add(12 * npart), add(12 * npart);
  return (b + 255) & ~(size_t)255;
}

static std::string plan_key(const aar_params* p, const aar_tables* t, int precision, int schedule) {
  std::string k;
  auto put = [&](const void* v, size_t n) { k.append((const char*)v, n); };
  const int i[] = {p->Nmesh, p->Nsample, p->Nmax, p->N_modes, p->SphereMode, p->FixedAmplitude, p->PhaseFlip,
                   p->WhichSpectrum, p->WhichTransfer, p->GlassTileFac, p->WDM_On, p->WDM_Vtherm_On, p->Variant,
                   precision, schedule, t->ntransfer, t->nglass};
  const double d[] = {p->Box, p->Omega, p->OmegaLambda, p->OmegaBaryon, p->OmegaDM_2ndSpecies, p->HubbleParam,
                      p->Sigma8, p->PrimordialIndex, p->Redshift, p->ShapeGamma, p->UnitLength_in_cm,
                      p->UnitMass_in_g, p->UnitVelocity_in_cm_per_s, p->InputSpectrum_UnitLength_in_cm,
                      p->WDM_PartMass_in_kev, p->A_res, p->w_res, p->phi_res, t->glass_box};
  put(i, sizeof i);
  put(d, sizeof d);
  return k;
}

static bool plan_matches(const Plan& pl, const std::string& key, const aar_params* p, const aar_tables* t) {
  const size_t nm = (size_t)p->N_modes;
  auto same = [](const auto& v, const auto* x, size_t n) {
    return v.size() == n && (n == 0 || memcmp(v.data(), x, sizeof(v[0]) * n) == 0);
  };
  const bool tab = p->WhichTransfer == 2 && t->ntransfer > 0 && t->transfer_k && t->transfer_t;
  return pl.key == key && same(pl.alpha, t->alpha, nm * nm * nm) && same(pl.x0in, t->x0, 2 * nm) &&
         (tab ? same(pl.tk, t->transfer_k, (size_t)t->ntransfer) && same(pl.tt, t->transfer_t, (size_t)t->ntransfer)
              : pl.tk.empty()) &&
         same(pl.glass_in, t->glass, 3 * (size_t)t->nglass);
}

// Device views of a plan's block and the kernels' k-space and transfer parameters.
static void plan_views(const aar_params* p, Plan& pl, aar_job* j) {
  char* b = (char*)pl.block;
  j->kv = (double*)(b + pl.o_kv);
  j->sinc = (double*)(b + pl.o_sinc);
  j->icic = (double*)(b + pl.o_icic);
  j->x0 = (double*)(b + pl.o_x0);
  j->outc = (unsigned char*)(b + pl.o_outc);
  j->modes = (int*)(b + pl.o_modes);
  j->kmodes = (int*)(b + pl.o_kmodes);
  j->groups = (int*)(b + pl.o_groups);
  j->terms = (NgTerm*)(b + pl.o_terms);
  j->weights = (NgWeight*)(b + pl.o_weights);
  j->fgroups = (int*)(b + pl.o_fgroups);
  j->frecs = (NgRec4*)(b + pl.o_frecs);
  j->logk = pl.nt ? (double*)(b + pl.o_logk) : nullptr;
  j->logT = pl.nt ? (double*)(b + pl.o_logT) : nullptr;
  j->glass = (float*)(b + pl.o_glass);
  j->nglass = pl.nglass;
  j->glass_box = pl.glass_box;
  j->c = pl.c;
  j->M = pl.M;
  j->K = pl.K;
  j->modes_h = pl.modes_h;
  j->kmodes_h = pl.kmodes_h;
  j->terms_h = pl.terms_h;
  j->groups_h = pl.groups_h;
  j->pairs = pl.pairs;
  KSpace& ks = j->ks;
  ks.kv = j->kv;
  ks.outc = j->outc;
  ks.sinc = j->sinc;
  ks.icic = j->icic;
  ks.sphere = p->SphereMode;
  ks.box = p->Box;
  ks.nsh = (double)(p->Nsample / 2);
  ks.s = (4 - p->PrimordialIndex) / 3;
  const double kmin = 2 * AAR_PI / p->Box, kmax = 2 * AAR_PI / p->Box * p->Nmax;
  ks.kmin_s = pow(kmin, (4 - p->PrimordialIndex) / 3);
  ks.kmax_s = pow(kmax, (4 - p->PrimordialIndex) / 3);
  ks.x0 = j->x0;
  ks.res = p->Variant != 1 ? 0 : (p->WhichSpectrum == 2 ? 2 : 3);
  ks.res_A = p->A_res;
  ks.res_w = p->w_res;
  ks.res_phi = p->phi_res;
  TransferDev& tf = j->tf;
  tf = TransferDev{};
  tf.which = p->WhichTransfer;
  if (p->WhichTransfer == 1) {  // k-independent part of tk_eh (power.c:526-540)
    const double hubble = p->HubbleParam, omegam = p->Omega;
    double ombh2 = p->OmegaBaryon * p->HubbleParam * p->HubbleParam;
    if (p->OmegaBaryon == 0) ombh2 = 0.04 * p->HubbleParam * p->HubbleParam;
    const double theta = 2.728 / 2.7, ommh2 = omegam * hubble * hubble;
    tf.kconv = 3.085678e24 / p->UnitLength_in_cm;
    tf.eh_s = 44.5 * log(9.83 / ommh2) / sqrt(1. + 10. * exp(0.75 * log(ombh2))) * hubble;
    tf.eh_a = 1. - 0.328 * log(431. * ommh2) * ombh2 / ommh2 +
              0.380 * log(22.3 * ommh2) * (ombh2 / ommh2) * (ombh2 / ommh2);
    tf.eh_theta2 = theta * theta;
    tf.eh_omh = omegam * hubble;
    tf.eh_twoe = 2. * exp(1.);
  } else if (p->WhichTransfer == 2) {
    tf.kconv = p->InputSpectrum_UnitLength_in_cm / p->UnitLength_in_cm;
    tf.nt = pl.nt;
    tf.logk = j->logk;
    tf.logT = j->logT;
  }
}

// The context's plan for (p, t, precision, schedule): found in its cache, or laid out,
// allocated as one block and uploaded.  At most kPlanCache plans stay cached; the least
// recently used plan that no job holds is dropped first.
static constexpr size_t kPlanCache = 32;

static int ctx_plan(aar_ctx* ctx, const aar_params* p, const aar_tables* t, int precision, int schedule,
                    std::shared_ptr<Plan>* out) {
  const std::string key = plan_key(p, t, precision, schedule);
  {
    std::lock_guard<std::mutex> g(ctx->plan_mu);
    for (auto& pl : ctx->plans)
      if (plan_matches(*pl, key, p, t)) {
        pl->last_use = ++ctx->plan_clock;
        *out = pl;
        return AAR_OK;
      }
  }
  auto pl = std::make_shared<Plan>();
  int rc = ctx_cosmology(ctx, p, t, &pl->c);
  if (!rc) rc = plan_layout(p, t, schedule, *pl);
  if (rc) return rc;
  pl->key = key;
  const size_t nm = (size_t)p->N_modes;
  pl->alpha.assign(t->alpha, t->alpha + nm * nm * nm);
  pl->x0in.assign(t->x0, t->x0 + 2 * nm);
  if (p->WhichTransfer == 2 && t->ntransfer > 0 && t->transfer_k && t->transfer_t) {
    pl->tk.assign(t->transfer_k, t->transfer_k + t->ntransfer);
    pl->tt.assign(t->transfer_t, t->transfer_t + t->ntransfer);
  }
  pl->glass_in.assign(t->glass, t->glass + 3 * (size_t)t->nglass);
  if ((rc = device_alloc(&pl->block, pl->bytes))) return rc;
  pl->owned = true;
  if (hipMemcpy(pl->block, pl->image.data(), pl->bytes, hipMemcpyHostToDevice) != hipSuccess) return AAR_ERR_HIP;
  std::vector<char>().swap(pl->image);  // the device holds it now
  std::lock_guard<std::mutex> g(ctx->plan_mu);
  for (auto& other : ctx->plans)  // built meanwhile by another thread: use that one
    if (plan_matches(*other, key, p, t)) {
      other->last_use = ++ctx->plan_clock;
      *out = other;
      return AAR_OK;
    }
  pl->last_use = ++ctx->plan_clock;
  ctx->plans.push_back(pl);
  while (ctx->plans.size() > kPlanCache) {
    size_t victim = ctx->plans.size();
    for (size_t q = 0; q < ctx->plans.size(); ++q)
      if (ctx->plans[q].use_count() == 1 &&
          (victim == ctx->plans.size() || ctx->plans[q]->last_use < ctx->plans[victim]->last_use))
        victim = q;
    if (victim == ctx->plans.size()) break;  // every cached plan is held by a job
    ctx->plans.erase(ctx->plans.begin() + victim);
  }
  *out = pl;
  return AAR_OK;
}

extern "C" long long aar_job_bytes(const aar_params* p, const aar_tables* t, int precision, int schedule) {
  if (validate(p, t)) return -1;
  Plan pl;
  if (plan_layout(p, t, schedule, pl)) return -1;
  return (long long)(pl.bytes + job_own_bytes(p, t, precision, schedule, pl.M, pl.K));
}

long long aar_job_ws_bytes(const aar_params* p, const aar_tables* t, int precision, int schedule, int plan_in_ws) {
  if (plan_in_ws) return aar_job_bytes(p, t, precision, schedule);
  if (validate(p, t)) return -1;
  const Schedule s = build_schedule(*p, t->alpha);
  return (long long)job_own_bytes(p, t, precision, schedule, (int)s.modes.size(), (int)s.kmodes.size());
}

extern "C" int aar_job_prepare(aar_ctx* ctx, const aar_params* p, const aar_tables* t, int precision,
                               int schedule, aar_job** out) {
  return aar_job_prepare_ws(ctx, p, t, precision, schedule, nullptr, 0, nullptr, 0, out, nullptr, nullptr);
}

// Prepare a job.  ws = nullptr: the job's own buffers are hipMalloc'd; otherwise they are carved
// from `ws` (ws_size bytes, at least aar_job_ws_bytes) and the uploads are enqueued on `up`.
// plan_in_ws = 0: the tables come from the context's plan cache; 1: the job lays out its own
// tables at the start of `ws` (no allocation anywhere).
int aar_job_prepare_ws(aar_ctx* ctx, const aar_params* p, const aar_tables* t, int precision, int schedule,
                       void* ws, size_t ws_size, hipStream_t up, int plan_in_ws, aar_job** out, float* out_pos,
                       float* out_vel) {
  if (!ctx || !out) return AAR_ERR_BAD_ARG;
  int rc = validate(p, t);
  if (rc) return rc;
  if (precision != AAR_FP32 && precision != AAR_FP64) return AAR_ERR_BAD_ARG;
  if (schedule != AAR_SCHED_PLAIN && schedule != AAR_SCHED_BATCHED) return AAR_ERR_BAD_ARG;
  if (plan_in_ws && !ws) return AAR_ERR_BAD_ARG;
  AAR_HIP(hipSetDevice(ctx->device));
  aar_job* j = new aar_job();
  j->ctx = ctx;
  j->p = *p;
  j->prec = precision;
  j->sched = schedule;
  j->cxsize = precision == AAR_FP32 ? 8 : 16;
  j->g = make_geom(p->Nmesh);
  j->ws = (char*)ws;
  j->ws_size = ws_size;
  j->up = up;
  auto fail = [&](int code) {
    aar_job_destroy(j);
    return code;
  };
  const int N = p->Nmesh;
  if (plan_in_ws) {
    auto pl = std::make_shared<Plan>();
    if ((rc = ctx_cosmology(ctx, p, t, &pl->c)) || (rc = plan_layout(p, t, schedule, *pl))) return fail(rc);
    if ((rc = job_alloc(j, &pl->block, pl->bytes))) return fail(rc);
    if (up) {  // the image stays alive with the plan until the job is destroyed
      if (hipMemcpyAsync(pl->block, pl->image.data(), pl->bytes, hipMemcpyHostToDevice, up) != hipSuccess)
        return fail(AAR_ERR_HIP);
    } else if (hipMemcpy(pl->block, pl->image.data(), pl->bytes, hipMemcpyHostToDevice) != hipSuccess) {
      return fail(AAR_ERR_HIP);
    }
    j->plan = pl;
  } else if ((rc = ctx_plan(ctx, p, t, precision, schedule, &j->plan))) {
    return fail(rc);
  }
  plan_views(p, *j->plan, j);
  // the job's own buffers
  j->nslots = slots_needed(j->M, j->K, schedule, N);
  if ((rc = job_alloc(j, &j->pool, (size_t)j->nslots * j->g.fs * j->cxsize))) return fail(rc);
  if (schedule == AAR_SCHED_PLAIN && (rc = job_alloc(j, (void**)&j->acc, (size_t)j->g.fs * 16)))
    return fail(rc);
  if (!ws && N >= 256 && is_pow2(N)) {  // first job of this size and precision: tune the x / y tiles on its pool
    bool tuned;
    {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      tuned = g_tuned[precision == AAR_FP64][ilog2i(N)];
    }
    const int nf = std::min(j->nslots, 8);
    if (!tuned && (rc = precision == AAR_FP32 ? autotune_t<float>(ctx, N, nf, j->pool, j->nslots, nullptr, nullptr)
                                              : autotune_t<double>(ctx, N, nf, j->pool, j->nslots, nullptr, nullptr)))
      return fail(rc);
  }
  std::vector<uint32_t> seeds((size_t)N * N);
  aar_seedtable(N, p->Seed, seeds.data());
  if ((rc = job_upload(j, &j->seeds, seeds))) return fail(rc);
  j->fused_lpt = schedule == AAR_SCHED_BATCHED && N >= 64 && is_pow2(N);
  {
    int steps = 0;  // recurrence steps of P_{n-1} per value over the mode fields (n >= 4)
    for (int n : j->plan->modes_h) steps += n >= 4 ? n - 2 : 0;
    j->short_filters = steps <= 2 * (int)j->plan->modes_h.size();
  }
  // the fused z pass of the NG sum (Exec::run's condition), compiled for this schedule; not in
  // the allocation-free path (loading a module allocates)
  if (ctx->jit && !plan_in_ws && schedule == AAR_SCHED_BATCHED && j->M >= 1 && j->M <= 16 && N >= 64 &&
      is_pow2(N)) {
    const Plan& pl = *j->plan;
    const std::string src = jit_zng_source(precision == AAR_FP64, N / 2, j->M, j->K, pl.fgroups_h.data(),
                                           pl.frecs_h.data());
    jit_zng_shape(precision == AAR_FP64, N / 2, j->M, &j->jit.rows, &j->jit.threads);
    std::string log;
    j->jit_status = jit_zng_kernel(ctx->device, src, &j->jit, &log);
    if (j->jit_status) {  // the library's k_z_ng runs instead
      j->jit.fn = nullptr;
      fprintf(stderr, "aarambam: fused z kernel for this schedule not compiled (status %d): %.400s\n",
              j->jit_status, log.c_str());
    }
  }
  j->s_disp = j->s_reg + 6;
  j->s_disp2 = j->s_reg + (j->fused_lpt ? 1 : 0);
  j->npart = (long long)p->GlassTileFac * p->GlassTileFac * p->GlassTileFac * t->nglass;
  if (out_pos && out_vel) {  // the caller's device arrays
    j->pos = out_pos;
    j->vel = out_vel;
    j->ext_out = true;
  } else if ((rc = job_alloc(j, (void**)&j->pos, sizeof(float) * 3 * j->npart)) ||
             (rc = job_alloc(j, (void**)&j->vel, sizeof(float) * 3 * j->npart))) {
    return fail(rc);
  }
  if (!j->ext_out) {
    AAR_HIP(hipEventCreate(&j->ev0));
    AAR_HIP(hipEventCreate(&j->ev1));
  }
  *out = j;
  return AAR_OK;
}

static void drop_graph(aar_job* j);

extern "C" int aar_job_reseed(aar_job* j, int seed, double fnl) {
  if (!j || j->ws) return AAR_ERR_BAD_ARG;
  AAR_HIP(hipSetDevice(j->ctx->device));
  const size_t n = (size_t)j->g.N * j->g.N;
  if (!j->stage[0]) {
    for (int b = 0; b < 2; ++b) {
      if (hipHostMalloc((void**)&j->stage[b], n * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess) {
        j->stage[b] = nullptr;
        return AAR_ERR_ALLOC;
      }
      AAR_HIP(hipEventCreateWithFlags(&j->stage_ev[b], hipEventDisableTiming));
    }
  }
  const int b = j->stage_next;
  if (j->stage_pending == b) return AAR_ERR_BAD_ARG;  // cannot happen: two buffers alternate
  if (j->stage_used[b]) AAR_HIP(hipEventSynchronize(j->stage_ev[b]));  // its last copy has finished
  aar_seedtable(j->g.N, seed, j->stage[b]);
  j->stage_pending = b;
  j->stage_next = b ^ 1;
  j->p.Seed = seed;
  if (fnl != j->p.Fnl) {
    j->p.Fnl = fnl;
    if (j->graph_exec) {  // fNL is a kernel argument of the captured graph: capture again
      if (j->ran) AAR_HIP(hipEventSynchronize(j->ev1));
      drop_graph(j);
    }
  }
  return AAR_OK;
}

extern "C" int aar_job_run_noalloc(aar_ctx* ctx, const aar_params* p, const aar_tables* t, int precision,
                                   int schedule, float* pos, float* vel, void* stream) {
  if (!ctx || !pos || !vel) return AAR_ERR_BAD_ARG;
  int rc = validate(p, t);
  if (rc) return rc;
  if (precision != AAR_FP32 && precision != AAR_FP64) return AAR_ERR_BAD_ARG;
  if (schedule != AAR_SCHED_PLAIN && schedule != AAR_SCHED_BATCHED) return AAR_ERR_BAD_ARG;
  {
    Plan pl;
    if ((rc = plan_layout(p, t, schedule, pl))) return rc;
    if (pl.bytes + job_own_bytes(p, t, precision, schedule, pl.M, pl.K, false) > kScratchBytes)
      return AAR_ERR_SCRATCH;
  }
  void* ws = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    ws = g_scratch_ptr[ctx->device];
  }
  if (!ws) return AAR_ERR_HIP;  // set up by aar_ctx_create
  AAR_HIP(hipSetDevice(ctx->device));
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  aar_job* j = nullptr;
  if ((rc = aar_job_prepare_ws(ctx, p, t, precision, schedule, ws, kScratchBytes, st, 1, &j, pos, vel)))
    return rc;
  rc = precision == AAR_FP32 ? Exec<float>::run(j, st) : Exec<double>::run(j, st);
  // the uploads read host copies the job holds: wait for the work before dropping them
  if (hipStreamSynchronize(st) != hipSuccess && !rc) rc = AAR_ERR_HIP;
  aar_job_destroy(j);
  return rc;
}

extern "C" int aar_ctx_cached_plans(aar_ctx* ctx, int* count, long long* bytes) {
  if (!ctx) return AAR_ERR_BAD_ARG;
  std::lock_guard<std::mutex> g(ctx->plan_mu);
  long long b = 0;
  for (auto& pl : ctx->plans) b += (long long)pl->bytes;
  if (count) *count = (int)ctx->plans.size();
  if (bytes) *bytes = b;
  return AAR_OK;
}

extern "C" int aar_job_run(aar_job* j, void* stream) {
  if (!j) return AAR_ERR_BAD_ARG;
  AAR_HIP(hipSetDevice(j->ctx->device));
  hipStream_t st = stream ? (hipStream_t)stream : j->ctx->stream;
  if (j->stage_pending >= 0) {  // the seed table of aar_job_reseed, ahead of the job's kernels
    const int b = j->stage_pending;
    AAR_HIP(hipMemcpyAsync(j->seeds, j->stage[b], sizeof(uint32_t) * j->g.N * j->g.N, hipMemcpyHostToDevice, st));
    AAR_HIP(hipEventRecord(j->stage_ev[b], st));
    j->stage_used[b] = true;
    j->stage_pending = -1;
  }
  AAR_HIP(hipEventRecord(j->ev0, st));
  if (j->use_graph) {
    if (!j->graph_exec) {  // capture every launch of the job once, then replay
      // Per-kind event timing stays outside graphs: HIP cannot time events
      // recorded by captured event nodes.
      const bool prof = j->profiling;
      j->profiling = false;
      AAR_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      int rc = j->prec == AAR_FP32 ? Exec<float>::run(j, st) : Exec<double>::run(j, st);
      hipGraph_t gr = nullptr;
      hipError_t e = hipStreamEndCapture(st, &gr);
      j->profiling = prof;
      if (rc) {
        if (gr) hipGraphDestroy(gr);
        return rc;
      }
      if (e != hipSuccess) return AAR_ERR_HIP;
      j->graph = gr;
      AAR_HIP(hipGraphInstantiate(&j->graph_exec, gr, nullptr, nullptr, 0));
    }
    AAR_HIP(hipGraphLaunch(j->graph_exec, st));
  } else {
    // per-kind events accumulate over runs until aar_job_sync reads them
    int rc = j->prec == AAR_FP32 ? Exec<float>::run(j, st) : Exec<double>::run(j, st);
    if (rc) return rc;
  }
  AAR_HIP(hipEventRecord(j->ev1, st));
  j->last_stream = st;
  j->ran = true;
  return AAR_OK;
}

extern "C" int aar_job_sync(aar_job* j) {
  if (!j) return AAR_ERR_BAD_ARG;
  if (!j->ran) return AAR_OK;  // nothing enqueued yet
  AAR_HIP(hipSetDevice(j->ctx->device));
  AAR_HIP(hipEventSynchronize(j->ev1));
  float ms = 0;
  AAR_HIP(hipEventElapsedTime(&ms, j->ev0, j->ev1));
  j->last_seconds = ms * 1e-3;
  for (const auto& r : j->recs) {
    float e = 0;
    AAR_HIP(hipEventElapsedTime(&e, r.a, r.b));
    j->k_launch[r.kind] += r.dispatches;
    j->k_sec[r.kind] += e * 1e-3;
    j->k_bytes[r.kind] += r.bytes;
  }
  j->recs.clear();
  j->evnext = 0;
  return AAR_OK;
}

static void drop_graph(aar_job* j) {
  if (j->graph_exec) hipGraphExecDestroy(j->graph_exec);
  if (j->graph) hipGraphDestroy(j->graph);
  j->graph_exec = nullptr;
  j->graph = nullptr;
}

extern "C" int aar_job_set_graph(aar_job* j, int on) {
  if (!j) return AAR_ERR_BAD_ARG;
  if (j->ran) hipEventSynchronize(j->ev1);
  drop_graph(j);  // re-captured on the next run
  j->use_graph = on != 0;
  return AAR_OK;
}

extern "C" int aar_job_set_profiling(aar_job* j, int on) {
  if (!j) return AAR_ERR_BAD_ARG;
  j->profiling = on != 0;
  j->recs.clear();  // launches not yet read by aar_job_sync are not counted
  j->evnext = 0;
  for (int k = 0; k < AAR_NKINDS; ++k) {
    j->k_launch[k] = 0;
    j->k_sec[k] = 0;
    j->k_bytes[k] = 0;
  }
  return AAR_OK;
}

extern "C" int aar_ctx_set_jit(aar_ctx* ctx, int on) {
  if (!ctx) return AAR_ERR_BAD_ARG;
  ctx->jit = on != 0;
  return AAR_OK;
}

extern "C" int aar_ctx_set_modes_x(aar_ctx* ctx, int mode) {
  if (!ctx || mode < 0 || mode > 2) return AAR_ERR_BAD_ARG;
  ctx->modes_x = mode;
  return AAR_OK;
}

extern "C" int aar_job_jit(const aar_job* j) {
  if (!j) return AAR_ERR_BAD_ARG;
  return j->jit.fn ? 1 : j->jit_status;
}

extern "C" int aar_job_kernel_stats(aar_job* j, int kind, int* launches, double* seconds, double* bytes) {
  if (!j || kind < 0 || kind >= AAR_NKINDS) return AAR_ERR_BAD_ARG;
  if (launches) *launches = j->k_launch[kind];
  if (seconds) *seconds = j->k_sec[kind];
  if (bytes) *bytes = j->k_bytes[kind];
  return AAR_OK;
}

extern "C" double aar_job_last_seconds(aar_job* j) { return j ? j->last_seconds : -1; }

extern "C" int aar_job_device_output(aar_job* j, float** pos, float** vel, long long* npart) {
  if (!j) return AAR_ERR_BAD_ARG;
  if (pos) *pos = j->pos;
  if (vel) *vel = j->vel;
  if (npart) *npart = j->npart;
  return AAR_OK;
}

extern "C" int aar_job_fetch(aar_job* j, float* pos, float* vel) {
  if (!j || !j->ran) return AAR_ERR_BAD_ARG;
  AAR_HIP(hipSetDevice(j->ctx->device));
  AAR_HIP(hipEventSynchronize(j->ev1));
  if (pos) AAR_HIP(hipMemcpy(pos, j->pos, sizeof(float) * 3 * j->npart, hipMemcpyDeviceToHost));
  if (vel) AAR_HIP(hipMemcpy(vel, j->vel, sizeof(float) * 3 * j->npart, hipMemcpyDeviceToHost));
  return AAR_OK;
}

template <typename T> static int fetch_field_t(aar_job* j, int which, double* out) {
  using C2 = cx_t<T>;
  const Geom& g = j->g;
  const int N = g.N;
  std::vector<T> host((size_t)2 * g.fs);
  const C2* src = nullptr;
  C2* tmp = nullptr;
  if (which == AAR_FIELD_PHIG_SPEC || which == AAR_FIELD_PHI_SPEC) {
    src = (const C2*)j->slot(which == AAR_FIELD_PHIG_SPEC ? j->s_phig : j->s_phi);
  } else if (which >= AAR_FIELD_DISP_X && which <= AAR_FIELD_DISP2_Z) {
    const int a = which - AAR_FIELD_DISP_X;  // 0..5
    src = (const C2*)j->slot(a < 3 ? j->s_disp + a : j->s_disp2 + a - 3);
  } else if (which == AAR_FIELD_PHIG_REAL || which == AAR_FIELD_PHING_REAL) {
    if (device_alloc((void**)&tmp, (size_t)g.fs * sizeof(C2))) return AAR_ERR_ALLOC;
    const void* s = j->slot(which == AAR_FIELD_PHIG_REAL ? j->s_phig : j->s_phi);
    // stream-ordered: the context stream does not synchronise with the null stream
    AAR_HIP(hipMemcpyAsync(tmp, s, (size_t)g.fs * sizeof(C2), hipMemcpyDeviceToDevice, j->ctx->stream));
    const C2* tw = sizeof(T) == 4 ? (const C2*)j->ctx->tw32 : (const C2*)j->ctx->tw64;
    int rc = Pipeline<T>::c2r(tmp, 1, g, tw, j->ctx->stream);
    if (rc) return rc;
    AAR_HIP(hipStreamSynchronize(j->ctx->stream));
    src = tmp;
  } else {
    return AAR_ERR_BAD_ARG;
  }
  AAR_HIP(hipMemcpy(host.data(), src, (size_t)g.fs * sizeof(C2), hipMemcpyDeviceToHost));
  if (tmp) hipFree(tmp);
  if (which <= AAR_FIELD_PHI_SPEC) from_internal_spec<T>(host.data(), out, N);
  else from_internal_real<T>(host.data(), out, N);
  return AAR_OK;
}

// Sub-box of a real field: the rows of one x-block (8 consecutive x) for
// consecutive y are contiguous in the x-blocked layout, so each x-block of
// the box is one copy.
template <typename T>
static int fetch_box_t(aar_job* j, int which, int x0, int y0, int z0, int nx, int ny, int nz, double* out) {
  using C2 = cx_t<T>;
  const Geom& g = j->g;
  const int N = g.N;
  const T* src = nullptr;
  C2* tmp = nullptr;
  if (which >= AAR_FIELD_DISP_X && which <= AAR_FIELD_DISP2_Z) {
    const int a = which - AAR_FIELD_DISP_X;
    src = (const T*)j->slot(a < 3 ? j->s_disp + a : j->s_disp2 + a - 3);
  } else {
    if (device_alloc((void**)&tmp, (size_t)g.fs * sizeof(C2))) return AAR_ERR_ALLOC;
    const void* s = j->slot(which == AAR_FIELD_PHIG_REAL ? j->s_phig : j->s_phi);
    AAR_HIP(hipMemcpyAsync(tmp, s, (size_t)g.fs * sizeof(C2), hipMemcpyDeviceToDevice, j->ctx->stream));
    const C2* tw = sizeof(T) == 4 ? (const C2*)j->ctx->tw32 : (const C2*)j->ctx->tw64;
    int rc = Pipeline<T>::c2r(tmp, 1, g, tw, j->ctx->stream);
    if (rc) {
      hipFree(tmp);
      return rc;
    }
    AAR_HIP(hipStreamSynchronize(j->ctx->stream));
    src = (const T*)tmp;
  }
  std::vector<T> buf((size_t)ny * kXB * N);
  int rc = AAR_OK;
  for (int xb = x0 / kXB; xb <= (x0 + nx - 1) / kXB && !rc; ++xb) {
    const long long r0 = ((long long)xb * N + y0) * kXB;
    if (hipMemcpy(buf.data(), src + r0 * N, buf.size() * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess) {
      rc = AAR_ERR_HIP;
      break;
    }
    for (int x = std::max(x0, xb * kXB); x < std::min(x0 + nx, xb * kXB + kXB); ++x)
      for (int y = y0; y < y0 + ny; ++y) {
        const T* row = buf.data() + ((size_t)(y - y0) * kXB + (x & (kXB - 1))) * N + z0;
        double* o = out + ((size_t)(x - x0) * ny + (y - y0)) * nz;
        for (int z = 0; z < nz; ++z) o[z] = row[z];
      }
  }
  if (tmp) hipFree(tmp);
  return rc;
}

extern "C" int aar_job_fetch_field_box(aar_job* j, int which, int x0, int y0, int z0, int nx, int ny, int nz,
                                       double* out) {
  if (!j || !out || !j->ran || which < AAR_FIELD_DISP_X || which > AAR_FIELD_PHING_REAL) return AAR_ERR_BAD_ARG;
  const int N = j->g.N;
  if (x0 < 0 || y0 < 0 || z0 < 0 || nx < 1 || ny < 1 || nz < 1 || x0 + nx > N || y0 + ny > N || z0 + nz > N)
    return AAR_ERR_BAD_ARG;
  AAR_HIP(hipSetDevice(j->ctx->device));
  AAR_HIP(hipEventSynchronize(j->ev1));
  return j->prec == AAR_FP32 ? fetch_box_t<float>(j, which, x0, y0, z0, nx, ny, nz, out)
                             : fetch_box_t<double>(j, which, x0, y0, z0, nx, ny, nz, out);
}

extern "C" int aar_job_fetch_field(aar_job* j, int which, double* out) {
  if (!j || !out || !j->ran) return AAR_ERR_BAD_ARG;
  AAR_HIP(hipSetDevice(j->ctx->device));
  AAR_HIP(hipEventSynchronize(j->ev1));
  return j->prec == AAR_FP32 ? fetch_field_t<float>(j, which, out) : fetch_field_t<double>(j, which, out);
}

extern "C" int aar_job_destroy(aar_job* j) {
  if (!j) return AAR_OK;
  hipSetDevice(j->ctx->device);
  if (j->ran) hipEventSynchronize(j->ev1);
  if (!j->ws) {  // the job's own buffers; the tables belong to the plan
    void* bufs[] = {j->pool, j->seeds, j->acc, j->ext_out ? nullptr : j->pos, j->ext_out ? nullptr : j->vel};
    for (void* b : bufs)
      if (b) hipFree(b);
  } else if (j->up && !j->ran) {
    hipStreamSynchronize(j->up);  // uploads from the staged host copies may still be in flight
  }
  for (int b = 0; b < 2; ++b) {
    if (j->stage_ev[b]) {
      hipEventSynchronize(j->stage_ev[b]);
      hipEventDestroy(j->stage_ev[b]);
    }
    if (j->stage[b]) hipHostFree(j->stage[b]);
  }
  j->plan.reset();
  if (j->graph_exec) hipGraphExecDestroy(j->graph_exec);
  if (j->graph) hipGraphDestroy(j->graph);
  for (hipEvent_t e : j->evpool) hipEventDestroy(e);
  if (j->ev0) hipEventDestroy(j->ev0);
  if (j->ev1) hipEventDestroy(j->ev1);
  delete j;
  return AAR_OK;
}

// ------------------------------------------------------------------ template check
// P(k) of phi_G and the cross bispectrum <phi_G phi_G psi> of the non-Gaussian
// part psi = (Phi - phi_G) N^3 / Fnl, with their expectations from the input
// P(k) and modal table.  With P(k) = E|phi_G(k)|^2 and Phi = phi_G + Fnl/N^3 psi,
// psi(k) = sum_K g_K(k) R2C[sum_t c_t f_i f_j], f_i = C2R[q_i phi_G]:
//   E[phi(k1) phi(k2) psi(k3)] = N^3 P(k1) P(k2) sum_K g_K(k3)
//                                 sum_{t in K} c_t (q_i(k1) q_j(k2) + q_j(k1) q_i(k2)),
// a separable sum, so its bin average is a product sum of filtered fields.
static void stat_reduce(const std::vector<double>& part, int nout, int nchunks, std::vector<double>& out) {
  out.assign(nout, 0.0);
  for (int o = 0; o < nout; ++o)
    for (int c = 0; c < nchunks; ++c) out[o] += part[(size_t)o * nchunks + c];
}

template <typename T>
static int ng_statistics_t(aar_job* j, int nb, const double* edges_h, double* pk_meas, double* pk_theory,
                           double* kmean, double* nmodes, double* b_meas, double* b_theory, double* ntri) {
  using C2 = cx_t<T>;
  const Geom& g = j->g;
  const int M = (int)j->modes_h.size(), K = (int)j->kmodes_h.size();
  const double N3 = (double)g.N * g.N * g.N;
  // field ids: I_phi, I_psi, J (ones), P, |k|, Q_m (m < M), G_K (K < K): nb fields each
  const int fIp = 0, fIs = nb, fJ = 2 * nb, fP = 3 * nb, fK = 4 * nb, fQ = 5 * nb, fG = (5 + M) * nb;
  const int nf = (5 + M + K) * nb;
  double2* F = nullptr;
  double* ed = nullptr;
  int* ofs_d = nullptr;
  StatTerm* terms_d = nullptr;
  double* part_d = nullptr;
  int rc = device_alloc((void**)&F, (size_t)nf * g.fs * sizeof(double2));
  if (!rc) rc = device_alloc((void**)&ed, sizeof(double) * (nb + 1));
  hipStream_t st = j->ctx->stream;
  const C2* phig = (const C2*)j->slot(j->s_phig);
  const C2* phi = (const C2*)j->slot(j->s_phi);
  const double pamp = j->c.Anorm * j->c.fac * j->c.fac * j->c.Beta * j->c.Beta, ns = j->p.PrimordialIndex;
  if (!rc && hipMemcpy(ed, edges_h, sizeof(double) * (nb + 1), hipMemcpyHostToDevice) != hipSuccess) rc = AAR_ERR_HIP;
  auto bins = [&](int fid, const C2* a, const C2* b, double scale, int kind, int n) {
    if (!rc) rc = stat_bin_fields<T>(a, b, scale, kind, n, pamp, ns, F + (size_t)fid * g.fs, nb, ed, g, j->ks, st);
  };
  bins(fIp, phig, nullptr, 1.0, 0, 0);
  bins(fIs, phi, phig, j->p.Fnl != 0 ? N3 / j->p.Fnl : 0.0, 0, 0);
  bins(fJ, nullptr, nullptr, 0, 1, 0);
  bins(fP, nullptr, nullptr, 0, 2, -1);
  bins(fK, nullptr, nullptr, 0, 4, 0);
  for (int m = 0; m < M; ++m) bins(fQ + m * nb, nullptr, nullptr, 0, 2, j->modes_h[m]);
  for (int k = 0; k < K; ++k) bins(fG + k * nb, nullptr, nullptr, 0, 3, j->kmodes_h[k]);
  if (!rc) rc = Pipeline<double>::c2r(F, nf, g, j->ctx->tw64, st);
  // outputs: 4 pair sums per bin, then nb^3 measured, theory and triangle-count triples
  std::vector<int> ofs{0};
  std::vector<StatTerm> tl;
  auto out = [&](std::initializer_list<StatTerm> ts) {
    for (auto& t : ts) tl.push_back(t);
    ofs.push_back((int)tl.size());
  };
  for (int b = 0; b < nb; ++b) {
    out({{fIp + b, fIp + b, -1, 1.0}});  // N^3 sum |phi|^2
    out({{fJ + b, fP + b, -1, 1.0}});    // N^3 sum P
    out({{fJ + b, fK + b, -1, 1.0}});    // N^3 sum |k|
    out({{fJ + b, fJ + b, -1, 1.0}});    // N^3 n_b
  }
  for (int b1 = 0; b1 < nb; ++b1)
    for (int b2 = 0; b2 < nb; ++b2)
      for (int b3 = 0; b3 < nb; ++b3) {
        out({{fIp + b1, fIp + b2, fIs + b3, 1.0}});
        for (int kk = 0; kk < K; ++kk)
          for (int t = j->groups_h[kk]; t < j->groups_h[kk + 1]; ++t) {
            const NgTerm& tm = j->terms_h[t];
            tl.push_back({fQ + tm.i * nb + b1, fQ + tm.j * nb + b2, fG + kk * nb + b3, tm.c * N3});
            tl.push_back({fQ + tm.j * nb + b1, fQ + tm.i * nb + b2, fG + kk * nb + b3, tm.c * N3});
          }
        ofs.push_back((int)tl.size());
        out({{fJ + b1, fJ + b2, fJ + b3, 1.0 / N3}});
      }
  const int nout = (int)ofs.size() - 1, nchunks = 64;
  if (!rc) rc = upload(&ofs_d, ofs);
  if (!rc) rc = upload(&terms_d, tl);
  if (!rc) rc = device_alloc((void**)&part_d, sizeof(double) * (size_t)nout * nchunks);
  std::vector<double> part((size_t)nout * nchunks), sums;
  for (int o0 = 0; o0 < nout && !rc; o0 += 65535) {
    const int n = std::min(nout - o0, 65535);
    rc = stat_sums(F, g, ofs_d + o0, terms_d, n, nchunks, part_d + (size_t)o0 * nchunks, st);
  }
  if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = AAR_ERR_HIP;
  if (!rc && hipMemcpy(part.data(), part_d, sizeof(double) * part.size(), hipMemcpyDeviceToHost) != hipSuccess)
    rc = AAR_ERR_HIP;
  if (!rc) {
    stat_reduce(part, nout, nchunks, sums);
    for (int b = 0; b < nb; ++b) {
      const double n = sums[4 * b + 3] / N3;
      if (nmodes) nmodes[b] = n;
      if (pk_meas) pk_meas[b] = n > 0 ? sums[4 * b] / N3 / n : 0;
      if (pk_theory) pk_theory[b] = n > 0 ? sums[4 * b + 1] / N3 / n : 0;
      if (kmean) kmean[b] = n > 0 ? sums[4 * b + 2] / N3 / n : 0;
    }
    const size_t t0 = 4 * (size_t)nb, nt = (size_t)nb * nb * nb;
    for (size_t q = 0; q < nt; ++q) {
      if (b_meas) b_meas[q] = sums[t0 + 3 * q];
      if (b_theory) b_theory[q] = sums[t0 + 3 * q + 1];
      if (ntri) ntri[q] = sums[t0 + 3 * q + 2];
    }
  }
  void* bufs[] = {F, ed, ofs_d, terms_d, part_d};
  for (void* b : bufs)
    if (b) hipFree(b);
  return rc;
}

extern "C" int aar_job_ng_statistics(aar_job* j, int nb, const double* edges, double* pk_meas, double* pk_theory,
                                     double* kmean, double* nmodes, double* b_meas, double* b_theory, double* ntri) {
  if (!j || !j->ran || nb < 1 || nb > 32 || !edges) return AAR_ERR_BAD_ARG;
  for (int b = 0; b < nb; ++b)
    if (!(edges[b] < edges[b + 1])) return AAR_ERR_BAD_ARG;
  AAR_HIP(hipSetDevice(j->ctx->device));
  AAR_HIP(hipEventSynchronize(j->ev1));
  return j->prec == AAR_FP32
             ? ng_statistics_t<float>(j, nb, edges, pk_meas, pk_theory, kmean, nmodes, b_meas, b_theory, ntri)
             : ng_statistics_t<double>(j, nb, edges, pk_meas, pk_theory, kmean, nmodes, b_meas, b_theory, ntri);
}

// ------------------------------------------------------------------ building blocks
template <typename T> static int fft3d_t(aar_ctx* ctx, int N, int dir, const double* in, double* out) {
  using C2 = cx_t<T>;
  const Geom g = make_geom(N);
  std::vector<T> host((size_t)2 * g.fs, (T)0);
  const long long n3 = (long long)N * N * N;
  if (dir > 0) to_internal_spec<T>(in, host.data(), N);
  else to_internal_real<T>(in, host.data(), N);
  C2* d = nullptr;
  if (device_alloc((void**)&d, (size_t)g.fs * sizeof(C2))) return AAR_ERR_ALLOC;
  AAR_HIP(hipMemcpyAsync(d, host.data(), (size_t)g.fs * sizeof(C2), hipMemcpyHostToDevice, ctx->stream));
  const C2* tw = sizeof(T) == 4 ? (const C2*)ctx->tw32 : (const C2*)ctx->tw64;
  int rc = dir > 0 ? Pipeline<T>::c2r(d, 1, g, tw, ctx->stream) : Pipeline<T>::r2c(d, 1, g, tw, ctx->stream);
  if (rc) {
    hipFree(d);
    return rc;
  }
  AAR_HIP(hipStreamSynchronize(ctx->stream));
  AAR_HIP(hipMemcpy(host.data(), d, (size_t)g.fs * sizeof(C2), hipMemcpyDeviceToHost));
  hipFree(d);
  if (dir > 0) from_internal_real<T>(host.data(), out, N);
  else from_internal_spec<T>(host.data(), out, N);
  return AAR_OK;
}

extern "C" int aar_fft3d(aar_ctx* ctx, int n, int precision, int dir, const double* in, double* out) {
  if (!ctx || !in || !out || !mesh_ok(n) || (dir != 1 && dir != -1)) return AAR_ERR_BAD_ARG;
  AAR_HIP(hipSetDevice(ctx->device));
  return precision == AAR_FP32 ? fft3d_t<float>(ctx, n, dir, in, out) : fft3d_t<double>(ctx, n, dir, in, out);
}

extern "C" int aar_gaussian_field(aar_ctx* ctx, const aar_params* p, const aar_tables* t, int precision,
                                  double* out) {
  aar_job* j = nullptr;
  int rc = aar_job_prepare(ctx, p, t, precision, AAR_SCHED_BATCHED, &j);
  if (rc) return rc;
  const double phase_shift = p->PhaseFlip == 1 ? AAR_PI : 0.0;
  if (precision == AAR_FP32)
    rc = Pipeline<float>::gauss((float2*)j->slot(0), j->g, j->ks, j->seeds, j->c.Anorm, p->PrimordialIndex,
                                j->c.fac, j->c.Beta, phase_shift, p->FixedAmplitude, ctx->stream);
  else
    rc = Pipeline<double>::gauss((double2*)j->slot(0), j->g, j->ks, j->seeds, j->c.Anorm,
                                 p->PrimordialIndex, j->c.fac, j->c.Beta, phase_shift, p->FixedAmplitude,
                                 ctx->stream);
  if (!rc && hipEventRecord(j->ev1, ctx->stream) != hipSuccess) rc = AAR_ERR_HIP;
  if (!rc) {
    j->ran = true;
    rc = aar_job_fetch_field(j, AAR_FIELD_PHIG_SPEC, out);
  }
  aar_job_destroy(j);
  return rc;
}

// Microbenchmark of one building block on nf fields of an N^3 grid:
// kind 0 x-pass, 1 y-pass, 2 z c2r, 3 z r2c, 4 plain copy kernel over nf spectra, 5 / 6 the x / y
// pass's tile access without the transform (k_xy_pattern).
// Returns the mean seconds per launch over `iters` launches.
template <typename T> static int bench_pass_t(aar_ctx* ctx, int N, int kind, int nf, int iters, double* sec) {
  using C2 = cx_t<T>;
  const Geom g = make_geom(N);
  C2 *d = nullptr, *d2 = nullptr;
  if (device_alloc((void**)&d, (size_t)nf * g.fs * sizeof(C2))) return AAR_ERR_ALLOC;
  if (kind == 4 && device_alloc((void**)&d2, (size_t)nf * g.fs * sizeof(C2))) {
    hipFree(d);
    return AAR_ERR_ALLOC;
  }
  // non-zero data: HBM moves zero-filled buffers measurably faster than job data
  int rc = aar_noise_fill<T>((T*)d, (long long)nf * g.fs * 2, ctx->stream);
  const C2* tw = sizeof(T) == 4 ? (const C2*)ctx->tw32 : (const C2*)ctx->tw64;
  auto launch = [&]() -> int {
    switch (kind) {
      case 0: return Pipeline<T>::xy_pass(d, nf, g, tw, 0, +1, ctx->stream);
      case 1: return Pipeline<T>::xy_pass(d, nf, g, tw, 1, +1, ctx->stream);
      case 2: return Pipeline<T>::z_pass(d, nf, g, tw, +1, ctx->stream);
      case 3: return Pipeline<T>::z_pass(d, nf, g, tw, -1, ctx->stream);
      case 5: return aar_xy_pattern<T>(d, nf, g, 0, ctx->stream);
      case 6: return aar_xy_pattern<T>(d, nf, g, 1, ctx->stream);
      default:
        return aar_copy16(d, d2, (long long)nf * g.fs * (long long)sizeof(C2), ctx->stream);
    }
  };
  if (!rc) rc = launch();
  hipEvent_t a, b;
  hipEventCreate(&a);
  hipEventCreate(&b);
  hipEventRecord(a, ctx->stream);
  for (int q = 0; q < iters && !rc; ++q) rc = launch();
  hipEventRecord(b, ctx->stream);
  hipEventSynchronize(b);
  float ms = 0;
  hipEventElapsedTime(&ms, a, b);
  hipEventDestroy(a);
  hipEventDestroy(b);
  *sec = ms * 1e-3 / iters;
  hipFree(d);
  if (d2) hipFree(d2);
  return rc;
}

extern "C" int aar_bench_pass(aar_ctx* ctx, int n, int precision, int kind, int nf, int iters, double* sec) {
  if (!ctx || !sec || !pow2(n) || nf < 1 || iters < 1 || kind < 0 || kind > 6) return AAR_ERR_BAD_ARG;
  AAR_HIP(hipSetDevice(ctx->device));
  return precision == AAR_FP32 ? bench_pass_t<float>(ctx, n, kind, nf, iters, sec)
                               : bench_pass_t<double>(ctx, n, kind, nf, iters, sec);
}

extern "C" int aar_ranlxd_device(aar_ctx* ctx, uint32_t seed, int n, double* out) {
  if (!ctx || !out || n <= 0) return AAR_ERR_BAD_ARG;
  AAR_HIP(hipSetDevice(ctx->device));
  double* d = nullptr;
  if (device_alloc((void**)&d, sizeof(double) * n)) return AAR_ERR_ALLOC;
  int rc = aar_ranlxd_draws_dev(seed, n, d, ctx->stream);
  if (!rc) {
    AAR_HIP(hipStreamSynchronize(ctx->stream));
    AAR_HIP(hipMemcpy(out, d, sizeof(double) * n, hipMemcpyDeviceToHost));
  }
  hipFree(d);
  return rc;
}
