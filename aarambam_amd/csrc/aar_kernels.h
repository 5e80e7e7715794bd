// Kernel launchers of the NG-2LPT pipeline (implemented in aar_kernels.hip).
#pragma once
#include "aar_common.h"

// Per-job k-space tables, device pointers (built on the host with the
// reference's exact expressions, Aarambam/LPT/Basis/main.c:412-452).
struct KSpace {
  const double* kv;           // [N] kvec component of grid index i
  const unsigned char* outc;  // [N] 1 when |kv|*Box/(2 pi) > Nsample/2 (cube cut)
  const double* sinc;         // [N] CIC window sin(kv L/2N)/(kv L/2N), 1 at kv = 0
  const double* icic;         // [N] 1 / sinc^2, the CIC deconvolution of one axis
  int sphere;                 // SphereMode
  double box;                 // Box
  double nsh;                 // (double)(Nsample/2)
  double s;                   // (4 - ns)/3
  double kmin_s, kmax_s;      // kmin^s, kmax^s (Legendre argument map)
  const double* x0;           // [N_modes] X0Table[n][0]
  int res;                    // ResBasis oscillations of phi_G: bit 0 linear in k, bit 1 in ln k
  double res_A, res_w, res_phi;
};

// Transfer function parameters (power.c:32-63, 181-217, 519-548).
struct TransferDev {
  int which;           // 1 EH, 2 table, other: 1.0
  double kconv;        // unit conversion of k before evaluation
  double eh_s, eh_a, eh_theta2, eh_omh, eh_twoe;  // k-independent EH terms
  int nt;
  const double* logk;  // sorted log10 k
  const double* logT;  // log10 T - log10 T[0]
};

// One product term of the regrouped NG sum: h[K] += c * f[i] * f[j].
struct NgTerm {
  int i, j;  // field slots
  double c;
};

// The same sum factored for the fused z pass: each output group is a list of
// factors h[K] += f[v] * (sum_t c_t f[w_t]), one factor per mode slot of a
// vertex cover of the group's (i, j) pairs, so a pair costs one multiply-add
// and one register read instead of two of each.  A factor is a run of
// records, padded to a multiple of four (one 64-byte scalar load each): a
// header {v, n} and n parts {w_t, FP32 bits of c_t, c_t}.
struct NgRec {
  int slot;  // v (header) or w_t (part)
  int aux;   // n (header) or the FP32 coefficient's bits (part)
  double c;  // part: the coefficient
};
struct alignas(64) NgRec4 {
  NgRec q[4];
};

// Weight list entry: W(k) += c * g_n(k).
struct NgWeight {
  int n;     // mode index of the weight function
  double c;  // coefficient
};

// Part of a field an x / y pass covers: kz columns [k0, k0 + cols) of the kz < N/2 block
// (cols = 0: all of them) and, with nyq, the Nyquist plane.
struct XYRange {
  int k0 = 0, cols = 0;
  bool nyq = true;
};

// A multi-output y pass (inverse): the spectra that go into real space are products of one
// non-separable spectrum and per-axis factors, so fields that differ only in their factors of
// ky and kz share one x transform (and one field in HBM until the y pass).  The nin inputs sit
// in slots in0 .. in0 + nin - 1 after the nout outputs (slots 0 .. nout - 1); the x pass
// transforms them in place; the y pass transforms input src[o] once per output o, loaded times
// the factor code[o] of its (y, kz) grid indices:
//   bits 0-1 power of ky, bits 2-3 power of kz, bit 4 the CIC deconvolution of y and z
//   1 / (sinc_y^2 sinc_z^2), bit 5 times i, bit 6 negated.
struct XSplit {
  int nin = 0, nout = 0, in0 = 0;
  signed char src[9] = {};
  unsigned char code[9] = {};
  const double* kv = nullptr;    // [N] kvec component of a grid index (KSpace::kv)
  const double* icic = nullptr;  // [N] 1 / sinc^2 (KSpace::icic)
};
// The y pass of n outputs of a split in one launch: pair i writes slot dst[i] from slot src[i]
// times the factor code[i]; the pairs of one tile are adjacent workgroups.
struct XOut {
  int n = 0;
  signed char src[9] = {}, dst[9] = {};
  unsigned char code[9] = {};
  const double* kv = nullptr;
  const double* icic = nullptr;
};
// The x pass of the mode fields made from phi_G (modes_x): output m (slot m) is the inverse x
// transform of q_{modes[m]}(k) phi_G(k) (k_modes_spec's filter, cut and origin zero).
struct XModes {
  int M = 0;
  const void* phig = nullptr;
  const int* modes = nullptr;  // device, M mode numbers
  KSpace ks{};
};
enum : unsigned char { kXsKy = 1, kXsKy2 = 2, kXsKz = 4, kXsKz2 = 8, kXsCic = 16, kXsI = 32, kXsNeg = 64 };

template <typename T> struct Pipeline {
  // FFT passes on nf consecutive fields starting at f (stride g.fs)
  static int c2r(cx_t<T>* f, int nf, const Geom& g, const cx_t<T>* tw, hipStream_t st);
  static int r2c(cx_t<T>* f, int nf, const Geom& g, const cx_t<T>* tw, hipStream_t st);
  // variant < 0: the tile shape from the variant table (aar_set_xy_variant)
  // xs: the y pass (axis 1, dir > 0) of the xs->n outputs of a split (nf = 1); xm: the x pass
  // (axis 0, dir > 0) of the xm->M mode fields of phi_G into f (nf = 1)
  static int xy_pass(cx_t<T>* f, int nf, const Geom& g, const cx_t<T>* tw, int axis, int dir,
                     hipStream_t st, int variant = -1, XYRange r = XYRange(), const XOut* xs = nullptr,
                     const XModes* xm = nullptr);
  // x then y pass (dir > 0) or y then x (dir < 0).  group_cols > 0: field by field, in groups of
  // group_cols kz columns (the last group may be narrower), so the first pass's output of a group
  // can still be in the last-level cache when the second pass reads it; 0: each pass over all nf
  // fields at once; < 0: the tuned choice for this precision and N (aar_set_xy_group).
  static int xy_pair(cx_t<T>* f, int nf, const Geom& g, const cx_t<T>* tw, int dir, hipStream_t st,
                     int group_cols);
  // inverse x pass of xs's inputs (in place), then the multi-output y pass into the outputs
  // (XSplit, slots from f).  group_cols as in xy_pair (per input: its x pass, then the y pass of
  // its outputs, group by group).  Powers of two only.
  static int xy_split(cx_t<T>* f, const XSplit& xs, const Geom& g, const cx_t<T>* tw, hipStream_t st,
                      int group_cols);
  // the M mode fields of phi_G into real space along x and y: the x pass filters each tile of
  // phi_G by q_n on its loads (the mode fields are never stored in k-space), then the y passes;
  // group_cols as in xy_pair.  Powers of two only.
  static int modes_xy(const cx_t<T>* phig, cx_t<T>* out, int M, const int* modes, const Geom& g,
                      const KSpace& ks, const cx_t<T>* tw, hipStream_t st, int group_cols);
  static int z_pass(cx_t<T>* f, int nf, const Geom& g, const cx_t<T>* tw, int dir, hipStream_t st);

  // fused z pass of the NG sum: M mode fields (x, y already inverse transformed)
  // -> z C2R -> products -> z R2C -> K fields over the first K slots
  // groups: K + 1 offsets into recs, in records (multiples of four)
  static int z_ng(cx_t<T>* f, int M, int K, const int* groups, const NgRec4* recs, const Geom& g,
                  const cx_t<T>* tw, hipStream_t st);
  // fused z pass of the 2LPT source: the six gradient fields 00, 01, 02, 11, 12, 22 in
  // slots 0-5 (x, y already inverse transformed) -> z C2R -> source -> z R2C into slot 0
  static int z_source(cx_t<T>* f, const Geom& g, const cx_t<T>* tw, hipStream_t st);
  static int gauss(cx_t<T>* phi, const Geom& g, const KSpace& ks, const uint32_t* seeds, double anorm,
                   double ns, double fac, double beta, double phase_shift, int fixed_amp, hipStream_t st);
  // out[m] = q_{modes[m]}(k) * phi(k) for m < M (cut and origin -> 0)
  static int modes_spec(const cx_t<T>* phi, cx_t<T>* out, int M, const int* modes, const Geom& g,
                        const KSpace& ks, hipStream_t st);
  // real points: out[K] = sum_{terms of group K} c f[i] f[j]
  static int products(const cx_t<T>* fin, int M, cx_t<T>* fout, int K, const NgTerm* terms,
                      const int* group_start, const Geom& g, hipStream_t st);
  // acc(k) += W(k) * c(k), W from the weight list (pairwise path)
  static int ng_accumulate(double2* acc, const cx_t<T>* c, const NgWeight* w, int nw, const Geom& g,
                           const KSpace& ks, hipStream_t st);
  // Phi = phiG + fnl/N^3 * sum_K g_{Kmode[K]} hK   (batched path)
  static int ng_combine(const cx_t<T>* phig, const cx_t<T>* h, int K, const int* kmodes,
                        cx_t<T>* phi, double fnl, const Geom& g, const KSpace& ks, hipStream_t st);
  // Phi = phiG + fnl/N^3 * acc   (pairwise path)
  static int ng_combine_acc(const cx_t<T>* phig, const double2* acc, cx_t<T>* phi, double fnl,
                            double n3, const Geom& g, hipStream_t st);
  // 6 displacement-gradient spectra and 3 CIC-deconvolved ZA spectra
  static int lpt_spec(const cx_t<T>* phi, cx_t<T>* out9, const Geom& g, const KSpace& ks,
                      const TransferDev& tf, double dplus, double beta, hipStream_t st);
  static int source(cx_t<T>* dg6, const Geom& g, hipStream_t st);  // writes into dg6[3]
  // batched path: combine and the nine 2LPT spectra in one pass (out9 may alias h); split: only
  // the five x-pass inputs of lpt_xsplit() (slots 0-4 of out9), the rest of each spectrum's
  // factors applied by the multi-output y pass
  static int ng_lpt(const cx_t<T>* phig, const cx_t<T>* h, int K, const int* kmodes, cx_t<T>* phi,
                    cx_t<T>* out9, double fnl, const Geom& g, const KSpace& ks, const TransferDev& tf,
                    double dplus, double beta, hipStream_t st, bool split = false);
  // split: the two x-pass inputs of disp2_xsplit() (slots 0, 1 of out3)
  static int disp2_spec(const cx_t<T>* src, cx_t<T>* out3, const Geom& g, const KSpace& ks,
                        hipStream_t st, bool split = false);
  // disp3 / disp2_3: three consecutive real fields each (ZA, 2LPT)
  static int cic(const cx_t<T>* disp3, const cx_t<T>* disp2_3, const Geom& g, const float* glass,
                 int nglass, double gbox, int tile, double box, double vp1, double vp2, float* pos,
                 float* vel, long long npart, hipStream_t st);
};

// The splits of the nine 2LPT spectra (outputs: the gradients 00, 01, 02, 11, 12, 22 and the
// CIC-deconvolved ZA displacements x, y, z; inputs, ng_lpt's split outputs, in slots 9-13) and
// of the three 2LPT displacements (inputs, disp2_spec's split outputs, in slots 3-4).
XSplit lpt_xsplit(const KSpace& ks);
XSplit disp2_xsplit(const KSpace& ks);

// zero n doubles (graph-capturable; no memset nodes)
int aar_zero_fill(double* p, long long n, hipStream_t st);
// n values of a fixed pattern in [-1/2, 1/2) (autotuner data)
template <typename T> int aar_noise_fill(T* p, long long n, hipStream_t st);
int aar_copy16(const void* src, void* dst, long long bytes, hipStream_t st);
// read and write back nf fields with the x (axis 0) or y (axis 1) pass's tile addressing, no transform
template <typename T> int aar_xy_pattern(cx_t<T>* f, int nf, const Geom& g, int axis, hipStream_t st);

// ---- modal decomposition (aar_modal.hip): QQ = (w Q) Q^T and QS = S (w Q)^T over tetrapyd points
int modal_cov_launch(const double* w, const int* i1, const int* i2, const int* i3, int npts, const int* p1,
                     const int* p2, const int* p3, int nmodes, const double* evals, int pmax, int nk, int nsplit,
                     double* partial, double* cov, hipStream_t st);
int modal_qs_launch(const double* S, int nshapes, const double* w, const int* i1, const int* i2, const int* i3,
                    int npts, const int* p1, const int* p2, const int* p3, int nmodes, const double* evals, int pmax,
                    int nk, int nsplit, double* partial, double* qs, hipStream_t st);
// R = S - sum_j coef_j Q_{sel_j} at every point; partial[block] = sum of w R^2 over the block's points
// (ceil(npts / 256) blocks)
int modal_residual_launch(const double* S, const double* w, const int* i1, const int* i2, const int* i3, int npts,
                          const int* sel, int nsel, const double* coef, const int* p1, const int* p2, const int* p3,
                          const double* evals, int pmax, int nk, double* R, double* partial, hipStream_t st);
// ncand column-major (nsel+1)^2 matrices qq[idx][idx], idx = sel followed by cand[b]
int modal_bordered_launch(const double* qq, int nmodes, const int* sel, int nsel, const int* cand, int ncand,
                          double* out, hipStream_t st);
// min and max eigenvalue magnitude of nmat symmetric m x m matrices (overwritten): ext[2b], ext[2b + 1]
int modal_extremes_launch(double* mats, int m, int nmat, double* ext, hipStream_t st);

// ---- binned statistics (aar_stats.hip)
struct StatTerm {  // one weighted product of real fields a, b, c (c < 0: none)
  int a, b, c;
  double w;
};
// nb band-filtered FP64 spectra of one kind (see aar_stats.hip) at out + b fs
template <typename TS>
int stat_bin_fields(const cx_t<TS>* a, const cx_t<TS>* b, double scale, int kind, int n, double pamp, double ns,
                    double2* out, int nb, const double* edges, const Geom& g, const KSpace& ks, hipStream_t st);
// per-chunk partial sums of nout outputs, each a weighted sum of field products
int stat_sums(const double2* fields, const Geom& g, const int* ofs, const StatTerm* terms, int nout, int nchunks,
              double* partial, hipStream_t st);
// x / y pass tile-shape variants (0 .. kXYVariantCount-1) per precision, mesh size and axis
constexpr int kXYVariantCount = 4;
int aar_xy_variant(int fp64, int n, int axis);
void aar_set_xy_variant(int fp64, int n, int axis, int v);
int aar_xy_group(int fp64, int n);
// radix plans of the x / y passes (0 .. kXYPlanCount-1; others than 0 only for n = 512, 1024)
constexpr int kXYPlanCount = 3;
int aar_xy_plan(int fp64, int n);
void aar_set_xy_plan(int fp64, int n, int plan);
void aar_set_xy_group(int fp64, int n, int cols);
// column groups of the multi-output pass pairs (xy_split): the tuned choice, else the pairs'
int aar_xy_split_group(int fp64, int n);
void aar_set_xy_split_group(int fp64, int n, int cols);  // -1: follow the pairs' choice
// kernel launches of one xy_pair call (of one xy_split call: nf = its inputs)
int aar_xy_pair_dispatches(int nf, int H, int group_cols);
// n uniforms of one ranlxd1 stream on the device (test hook)
int aar_ranlxd_draws_dev(uint32_t seed, int n, double* out, hipStream_t st);

// Twiddle tables: tw[L + m] = exp(-2 pi i m / L) for every power of two L <= 2048.
template <typename T> void make_twiddles(cx_t<T>* host_out /* kTwTotal entries */);
