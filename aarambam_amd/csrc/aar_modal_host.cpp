// Host side of the modal decomposition on the device: the two contractions of
// Aarambam/basis/tetrapyd.c (compute_mode_bispectra_covariance, compute_QS)
// and the orthogonal-matching-pursuit expansion of Basis.basis_expansion
// (Aarambam/basis/cmbbest.pyx:320-490) built on them.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>


#include "../../include/aarambam.h"
#include "aar_common.h"
#include "aar_internal.h"
#include "aar_kernels.h"

// ------------------------------------------------------------------ modal decomposition
// Device versions of compute_mode_bispectra_covariance and compute_QS
// (Aarambam/basis/tetrapyd.c:111-208, 210-281), same arguments.  Every index
// is checked on the host first: the kernels gather from the mode table.
struct ModalBufs {
  double *w = nullptr, *ev = nullptr, *part = nullptr, *out = nullptr, *S = nullptr;
  int *i1 = nullptr, *i2 = nullptr, *i3 = nullptr, *p1 = nullptr, *p2 = nullptr, *p3 = nullptr;
  ~ModalBufs() {
    void* b[] = {w, ev, part, out, S, i1, i2, i3, p1, p2, p3};
    for (void* x : b)
      if (x) hipFree(x);
  }
};

static int modal_check(const double* w, const int* i1, const int* i2, const int* i3, int npts, const int* p1,
                       const int* p2, const int* p3, int nmodes, const double* evals, int pmax, int nk) {
  if (!w || !i1 || !i2 || !i3 || !p1 || !p2 || !p3 || !evals || npts < 1 || nmodes < 1 || pmax < 1 || nk < 1)
    return AAR_ERR_BAD_ARG;
  for (int t = 0; t < npts; ++t)
    if (i1[t] < 0 || i1[t] >= nk || i2[t] < 0 || i2[t] >= nk || i3[t] < 0 || i3[t] >= nk) return AAR_ERR_BAD_ARG;
  for (int n = 0; n < nmodes; ++n)
    if (p1[n] < 0 || p1[n] >= pmax || p2[n] < 0 || p2[n] >= pmax || p3[n] < 0 || p3[n] >= pmax)
      return AAR_ERR_BAD_ARG;
  return AAR_OK;
}

static int modal_upload(ModalBufs& b, const double* w, const int* i1, const int* i2, const int* i3, int npts,
                        const int* p1, const int* p2, const int* p3, int nmodes, const double* evals, int pmax,
                        int nk, hipStream_t st) {
  auto up = [&](void** d, const void* h, size_t n) -> int {
    if (aar_device_alloc(d, n)) return AAR_ERR_ALLOC;
    return hipMemcpyAsync(*d, h, n, hipMemcpyHostToDevice, st) == hipSuccess ? AAR_OK : AAR_ERR_HIP;
  };
  int rc;
  if ((rc = up((void**)&b.w, w, sizeof(double) * npts)) || (rc = up((void**)&b.i1, i1, sizeof(int) * npts)) ||
      (rc = up((void**)&b.i2, i2, sizeof(int) * npts)) || (rc = up((void**)&b.i3, i3, sizeof(int) * npts)) ||
      (rc = up((void**)&b.p1, p1, sizeof(int) * nmodes)) || (rc = up((void**)&b.p2, p2, sizeof(int) * nmodes)) ||
      (rc = up((void**)&b.p3, p3, sizeof(int) * nmodes)) ||
      (rc = up((void**)&b.ev, evals, sizeof(double) * (size_t)pmax * nk)))
    return rc;
  return AAR_OK;
}

extern "C" int aar_mode_bispectra_covariance(aar_ctx* ctx, double* cov, const double* w, const int* i1,
                                             const int* i2, const int* i3, int npts, const int* p1, const int* p2,
                                             const int* p3, int nmodes, const double* evals, int pmax, int nk) {
  if (!ctx || !cov) return AAR_ERR_BAD_ARG;
  int rc = modal_check(w, i1, i2, i3, npts, p1, p2, p3, nmodes, evals, pmax, nk);
  if (rc) return rc;
  AAR_HIP(hipSetDevice(aar_ctx_device(ctx)));
  hipStream_t st = aar_ctx_stream(ctx);
  ModalBufs b;
  if ((rc = modal_upload(b, w, i1, i2, i3, npts, p1, p2, p3, nmodes, evals, pmax, nk, st))) return rc;
  // split the points so that the launch has about 2048 blocks, at least 256 points each
  const int ntiles = (nmodes + 127) / 128, npairs = ntiles * (ntiles + 1) / 2;
  int nsplit = std::max(1, std::min((2048 + npairs - 1) / npairs, 64));
  nsplit = std::max(1, std::min(nsplit, npts / 256));
  if (aar_device_alloc((void**)&b.part, sizeof(double) * (size_t)nsplit * nmodes * nmodes) ||
      aar_device_alloc((void**)&b.out, sizeof(double) * (size_t)nmodes * nmodes))
    return AAR_ERR_ALLOC;
  if ((rc = modal_cov_launch(b.w, b.i1, b.i2, b.i3, npts, b.p1, b.p2, b.p3, nmodes, b.ev, pmax, nk, nsplit, b.part,
                             b.out, st)))
    return rc;
  AAR_HIP(hipMemcpyAsync(cov, b.out, sizeof(double) * (size_t)nmodes * nmodes, hipMemcpyDeviceToHost, st));
  AAR_HIP(hipStreamSynchronize(st));
  return AAR_OK;
}

extern "C" int aar_mode_QS(aar_ctx* ctx, double* QS, const double* S, int nshapes, const double* w, const int* i1,
                           const int* i2, const int* i3, int npts, const int* p1, const int* p2, const int* p3,
                           int nmodes, const double* evals, int pmax, int nk) {
  if (!ctx || !QS || !S || nshapes < 1) return AAR_ERR_BAD_ARG;
  int rc = modal_check(w, i1, i2, i3, npts, p1, p2, p3, nmodes, evals, pmax, nk);
  if (rc) return rc;
  AAR_HIP(hipSetDevice(aar_ctx_device(ctx)));
  hipStream_t st = aar_ctx_stream(ctx);
  ModalBufs b;
  if ((rc = modal_upload(b, w, i1, i2, i3, npts, p1, p2, p3, nmodes, evals, pmax, nk, st))) return rc;
  const int nsplit = std::max(1, std::min(64, npts / 4096));
  if (aar_device_alloc((void**)&b.S, sizeof(double) * (size_t)nshapes * npts) ||
      aar_device_alloc((void**)&b.part, sizeof(double) * (size_t)nsplit * nshapes * nmodes) ||
      aar_device_alloc((void**)&b.out, sizeof(double) * (size_t)nshapes * nmodes))
    return AAR_ERR_ALLOC;
  AAR_HIP(hipMemcpyAsync(b.S, S, sizeof(double) * (size_t)nshapes * npts, hipMemcpyHostToDevice, st));
  if ((rc = modal_qs_launch(b.S, nshapes, b.w, b.i1, b.i2, b.i3, npts, b.p1, b.p2, b.p3, nmodes, b.ev, pmax, nk,
                            nsplit, b.part, b.out, st)))
    return rc;
  AAR_HIP(hipMemcpyAsync(QS, b.out, sizeof(double) * (size_t)nshapes * nmodes, hipMemcpyDeviceToHost, st));
  AAR_HIP(hipStreamSynchronize(st));
  return AAR_OK;
}

// ------------------------------------------------------------------ basis expansion
// Basis.basis_expansion (Aarambam/basis/cmbbest.pyx:320-490): expand shapes S
// (evaluated at the tetrapyd points) in the normalised mode bispectra by
// orthogonal matching pursuit, each step a conjugate-gradient solve of the
// normal equations of the selected modes.
namespace {

// scipy.sparse.linalg.cg (x0 = 0, M = I, atol = 0): stop when |r| < rtol |b| at
// the start of an iteration, at most maxiter iterations.
void cg_solve(const std::vector<double>& A, int n, const std::vector<double>& b, double rtol, long long maxiter,
              std::vector<double>& x) {
  x.assign(n, 0.0);
  double bn = 0;
  for (double v : b) bn += v * v;
  bn = sqrt(bn);
  if (bn == 0) return;
  const double tol = rtol * bn;
  std::vector<double> r = b, p(n), q(n);
  double rho_prev = 0;
  for (long long it = 0; it < maxiter; ++it) {
    double rn = 0;
    for (double v : r) rn += v * v;
    if (sqrt(rn) < tol) return;
    const double rho = rn;  // r . z with z = r
    if (it > 0) {
      const double beta = rho / rho_prev;
      for (int i = 0; i < n; ++i) p[i] = p[i] * beta + r[i];
    } else {
      p = r;
    }
    // q = A p, eight rows at a time: eight independent sums, each still taken
    // in column order, so q is the same as row by row
    double pq = 0;
    int i = 0;
    for (; i + 8 <= n; i += 8) {
      double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      const double* row = &A[(size_t)i * n];
      for (int j = 0; j < n; ++j) {
        const double pj = p[j];
        for (int r = 0; r < 8; ++r) s[r] += row[(size_t)r * n + j] * pj;
      }
      for (int r = 0; r < 8; ++r) {
        q[i + r] = s[r];
        pq += p[i + r] * s[r];
      }
    }
    for (; i < n; ++i) {
      double s = 0;
      const double* row = &A[(size_t)i * n];
      for (int j = 0; j < n; ++j) s += row[j] * p[j];
      q[i] = s;
      pq += p[i] * s;
    }
    const double alpha = rho / pq;
    for (int i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
    }
    rho_prev = rho;
  }
}

struct Dev {  // device buffers of one expansion
  std::vector<void*> all;
  ~Dev() {
    for (void* p : all) hipFree(p);
  }
  template <typename V> int alloc(V** p, size_t n) {
    if (aar_device_alloc((void**)p, sizeof(V) * n)) return AAR_ERR_ALLOC;
    all.push_back(*p);
    return AAR_OK;
  }
};

// np.linalg.cond (2-norm) of a symmetric matrix from its extreme eigenvalue
// magnitudes (min |l|, max |l|)
double cond_of(const double* ext) { return ext[0] > 0 ? ext[1] / ext[0] : INFINITY; }

}  // namespace

extern "C" int aar_mode_basis_expansion(aar_ctx* ctx, double* alpha, double* shape_cov, double* corr, double* mse,
                                        uint8_t* selected, const double* S, int nshapes, const double* w,
                                        const int* i1, const int* i2, const int* i3, int npts, const int* p1,
                                        const int* p2, const int* p3, int nmodes, const double* evals, int pmax,
                                        int nk, const uint8_t* bad_modes, double regularization, long long niter,
                                        double rtol, double rr_tol, double cond_tol, int max_modes) {
  if (!ctx || !alpha || !S || nshapes < 1 || niter < 1 || max_modes < -1) return AAR_ERR_BAD_ARG;
  int rc = modal_check(w, i1, i2, i3, npts, p1, p2, p3, nmodes, evals, pmax, nk);
  if (rc) return rc;
  // LDS of the residual kernel with every mode selected (the QS kernel needs less)
  if (sizeof(double) * ((size_t)pmax * nk + nmodes + 256) + sizeof(int) * 3 * (size_t)(nmodes + 1) > 64 * 1024)
    return AAR_ERR_UNSUPPORTED;
  AAR_HIP(hipSetDevice(aar_ctx_device(ctx)));
  hipStream_t st = aar_ctx_stream(ctx);
  ModalBufs b;
  if ((rc = modal_upload(b, w, i1, i2, i3, npts, p1, p2, p3, nmodes, evals, pmax, nk, st))) return rc;
  Dev d;
  const int qs_split = std::max(1, std::min(64, npts / 4096));
  const int ntiles = (nmodes + 127) / 128, npairs = ntiles * (ntiles + 1) / 2;
  int cov_split = std::max(1, std::min((2048 + npairs - 1) / npairs, 64));
  cov_split = std::max(1, std::min(cov_split, npts / 256));
  const int nrb = (npts + 255) / 256;
  double *S_d, *R_d, *qs_d, *qq_d, *part_d, *rr_d, *coef_d, *qqt_d;
  int *sel_d, *cand_d;
  if ((rc = d.alloc(&S_d, (size_t)nshapes * npts)) || (rc = d.alloc(&R_d, (size_t)npts)) ||
      (rc = d.alloc(&qs_d, (size_t)nshapes * nmodes)) || (rc = d.alloc(&qq_d, (size_t)nmodes * nmodes)) ||
      (rc = d.alloc(&qqt_d, (size_t)nmodes * nmodes)) ||
      (rc = d.alloc(&part_d, std::max((size_t)cov_split * nmodes * nmodes, (size_t)qs_split * nshapes * nmodes))) ||
      (rc = d.alloc(&rr_d, (size_t)nrb)) || (rc = d.alloc(&coef_d, (size_t)nmodes)) ||
      (rc = d.alloc(&sel_d, (size_t)nmodes)) || (rc = d.alloc(&cand_d, (size_t)nmodes)))
    return rc;
  AAR_HIP(hipMemcpyAsync(S_d, S, sizeof(double) * nshapes * (size_t)npts, hipMemcpyHostToDevice, st));
  // QQ (cmbbest.pyx:95-98 via tetrapyd.c) and QS = (S w) Q^T (cmbbest.pyx:347-351)
  if ((rc = modal_cov_launch(b.w, b.i1, b.i2, b.i3, npts, b.p1, b.p2, b.p3, nmodes, b.ev, pmax, nk, cov_split, part_d,
                             qq_d, st)) ||
      (rc = modal_qs_launch(S_d, nshapes, b.w, b.i1, b.i2, b.i3, npts, b.p1, b.p2, b.p3, nmodes, b.ev, pmax, nk,
                            qs_split, part_d, qs_d, st)))
    return rc;
  std::vector<double> QQ((size_t)nmodes * nmodes), QS((size_t)nshapes * nmodes);
  AAR_HIP(hipMemcpyAsync(QQ.data(), qq_d, sizeof(double) * QQ.size(), hipMemcpyDeviceToHost, st));
  AAR_HIP(hipMemcpyAsync(QS.data(), qs_d, sizeof(double) * QS.size(), hipMemcpyDeviceToHost, st));
  AAR_HIP(hipStreamSynchronize(st));
  const int n = nmodes;
  std::vector<double> norms(n);
  for (int i = 0; i < n; ++i) norms[i] = sqrt(QQ[(size_t)i * n + i]);
  // shape covariance (S w) S^T (cmbbest.pyx:343)
  std::vector<double> scov((size_t)nshapes * nshapes);
  for (int a = 0; a < nshapes; ++a)
    for (int c = 0; c < nshapes; ++c) {
      double s = 0;
      for (int t = 0; t < npts; ++t) s += S[(size_t)a * npts + t] * w[t] * S[(size_t)c * npts + t];
      scov[(size_t)a * nshapes + c] = s;
    }
  // modes dropped from the basis (cmbbest.pyx:366-398): QQ rows and columns
  // zeroed except the diagonal, QS and Q zeroed
  std::vector<uint8_t> bad(n);
  for (int i = 0; i < n; ++i)
    bad[i] = ((p1[i] <= 2 && p2[i] == 0 && p3[i] == 0) || (p1[i] == 1 && p2[i] == 1 && p3[i] == 0) ||
              (bad_modes && bad_modes[i]))
                 ? 1
                 : 0;
  std::vector<double> QQt((size_t)n * n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      const double v = (bad[i] || bad[j]) && i != j ? 0.0 : QQ[(size_t)i * n + j];
      QQt[(size_t)i * n + j] = v / norms[i] / norms[j];
    }
  AAR_HIP(hipMemcpyAsync(qqt_d, QQt.data(), sizeof(double) * QQt.size(), hipMemcpyHostToDevice, st));
  const int mmax = max_modes == 0 ? n : max_modes;
  // ncand symmetric m x m matrices (column-major, in bt.mats) and their extreme
  // eigenvalue magnitudes (bt.ext); the buffers grow with the selection and are
  // freed on every exit
  struct Batch {
    double *mats = nullptr, *ext = nullptr;
    size_t cap_m = 0, cap_v = 0;
    ~Batch() { release(); }
    void release() {
      if (mats) hipFree(mats);
      if (ext) hipFree(ext);
      mats = ext = nullptr;
    }
    int ensure(int m, int ncand) {
      const size_t nm = (size_t)m * m * ncand, nv = 2 * (size_t)ncand;
      if (nm <= cap_m && nv <= cap_v) return AAR_OK;
      release();
      cap_m = std::max(nm, cap_m + cap_m / 2);
      cap_v = std::max(nv, cap_v);
      if (aar_device_alloc((void**)&mats, sizeof(double) * cap_m) ||
          aar_device_alloc((void**)&ext, sizeof(double) * cap_v))
        return AAR_ERR_ALLOC;
      return AAR_OK;
    }
  } bt;
  std::vector<double> x;
  for (int model = 0; model < nshapes; ++model) {
    std::vector<double> qst(n);
    for (int i = 0; i < n; ++i) qst[i] = (bad[i] ? 0.0 : QS[(size_t)model * n + i]) / norms[i];
    std::vector<uint8_t> msk(n);
    int nsel = 0;
    for (int i = 0; i < n; ++i) {
      msk[i] = (max_modes == -1 || (p1[i] <= 3 && p2[i] <= 3 && p3[i] <= 3)) ? 1 : 0;
      nsel += msk[i];
    }
    auto solve = [&](double tol, long long maxit) {
      std::vector<int> idx;
      for (int i = 0; i < n; ++i)
        if (msk[i]) idx.push_back(i);
      const int m = (int)idx.size();
      std::vector<double> A((size_t)m * m), rhs(m);
      for (int r = 0; r < m; ++r) {
        for (int c = 0; c < m; ++c) A[(size_t)r * m + c] = QQt[(size_t)idx[r] * n + idx[c]];
        A[(size_t)r * m + r] += regularization;
        rhs[r] = qst[idx[r]];
      }
      cg_solve(A, m, rhs, tol, maxit, x);
      return idx;
    };
    double ext_here[2] = {0, 0};
    bool have_here = false;
    while (true) {
      if (n <= nsel) break;
      std::vector<int> idx = solve(rtol, niter);
      const int m = (int)idx.size();
      // residual of the first shape (the reference broadcasts S - sum_j Q~_j a_j
      // over all shapes and reads entry [0, 0]), its weighted norm RR
      std::vector<double> coef(m);
      for (int j = 0; j < m; ++j) coef[j] = bad[idx[j]] ? 0.0 : x[j] / norms[idx[j]];
      AAR_HIP(hipMemcpyAsync(sel_d, idx.data(), sizeof(int) * m, hipMemcpyHostToDevice, st));
      AAR_HIP(hipMemcpyAsync(coef_d, coef.data(), sizeof(double) * m, hipMemcpyHostToDevice, st));
      if ((rc = modal_residual_launch(S_d, b.w, b.i1, b.i2, b.i3, npts, sel_d, m, coef_d, b.p1, b.p2, b.p3, b.ev,
                                      pmax, nk, R_d, rr_d, st)))
        return rc;
      // queued behind the residual, read after one wait: its weighted norm RR, the
      // condition number of the selection and the correlation of every mode with
      // the residual (QS of R); the last two are not needed when RR already
      // converged.  The selection's condition number is that of the bordered
      // matrix the previous step picked (the same modes, so the same spectrum);
      // on the first step, its first m - 1 indices bordered by the last.
      std::vector<int> cand;
      for (int i = 0; i < n; ++i)
        if (!msk[i]) cand.push_back(i);
      const int nc = (int)cand.size();
      if ((rc = bt.ensure(m + 1, std::max(1, std::min(nc, 10))))) return rc;
      if (!have_here && ((rc = modal_bordered_launch(qqt_d, n, sel_d, m - 1, sel_d + (m - 1), 1, bt.mats, st)) ||
                         (rc = modal_extremes_launch(bt.mats, m, 1, bt.ext, st))))
        return rc;
      if ((rc = modal_qs_launch(R_d, 1, b.w, b.i1, b.i2, b.i3, npts, b.p1, b.p2, b.p3, nmodes, b.ev, pmax, nk,
                                qs_split, part_d, qs_d, st)))
        return rc;
      std::vector<double> rr(nrb), QR(n);
      AAR_HIP(hipMemcpyAsync(rr.data(), rr_d, sizeof(double) * nrb, hipMemcpyDeviceToHost, st));
      if (!have_here) AAR_HIP(hipMemcpyAsync(ext_here, bt.ext, sizeof(ext_here), hipMemcpyDeviceToHost, st));
      AAR_HIP(hipMemcpyAsync(QR.data(), qs_d, sizeof(double) * n, hipMemcpyDeviceToHost, st));
      AAR_HIP(hipStreamSynchronize(st));
      double RR = 0;
      for (double v : rr) RR += v;
      if (fabs(RR / scov[0]) <= rr_tol) break;
      if (nsel == mmax - 1) break;
      const double cond_here = cond_of(ext_here);
      if (cond_here > cond_tol) break;
      std::vector<double> qr(nc);
      for (int c = 0; c < nc; ++c) qr[c] = bad[cand[c]] ? 0.0 : QR[cand[c]] / norms[cand[c]] / sqrt(RR);
      // the ten candidates most correlated with the residual (ascending |QR|), of them the one
      // whose addition raises the condition number of the selection least.  The reference
      // evaluates the condition number with every candidate and reads it for these ten only,
      // so the bordered matrices of the ten are the only ones decomposed here.
      std::vector<int> order(nc);
      for (int c = 0; c < nc; ++c) order[c] = c;
      std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return fabs(qr[a]) < fabs(qr[c]); });
      const int first = std::max(0, nc - 10), ntop = nc - first;
      std::vector<int> top(ntop);
      for (int k = 0; k < ntop; ++k) top[k] = cand[order[first + k]];
      AAR_HIP(hipMemcpyAsync(cand_d, top.data(), sizeof(int) * ntop, hipMemcpyHostToDevice, st));
      if ((rc = modal_bordered_launch(qqt_d, n, sel_d, m, cand_d, ntop, bt.mats, st)) ||
          (rc = modal_extremes_launch(bt.mats, m + 1, ntop, bt.ext, st)))
        return rc;
      std::vector<double> ext(2 * (size_t)ntop);
      AAR_HIP(hipMemcpyAsync(ext.data(), bt.ext, sizeof(double) * ext.size(), hipMemcpyDeviceToHost, st));
      AAR_HIP(hipStreamSynchronize(st));
      std::vector<double> fcond(ntop);
      for (int k = 0; k < ntop; ++k) fcond[k] = cond_of(ext.data() + 2 * k) / cond_here - 1;
      int bk = 0;
      for (int k = 1; k < ntop; ++k)
        if (fcond[k] < fcond[bk]) bk = k;
      const int best = order[first + bk];
      msk[cand[best]] = 1;
      ++nsel;
      ext_here[0] = ext[2 * bk];
      ext_here[1] = ext[2 * bk + 1];
      have_here = true;
    }
    std::vector<int> idx = solve(1e-16, 1000000);
    for (int i = 0; i < n; ++i) alpha[(size_t)model * n + i] = 0.0;
    for (size_t j = 0; j < idx.size(); ++j) alpha[(size_t)model * n + idx[j]] = x[j] / norms[idx[j]];
    if (selected)
      for (int i = 0; i < n; ++i) selected[(size_t)model * n + i] = msk[i];
  }
  // convergence of the expansion (cmbbest.pyx:478-486)
  for (int a = 0; a < nshapes; ++a) {
    const double* al = alpha + (size_t)a * n;
    double ss = scov[(size_t)a * nshapes + a], sr = 0, rr = 0;
    for (int i = 0; i < n; ++i) sr += al[i] * QS[(size_t)a * n + i];
    for (int i = 0; i < n; ++i) {
      double s = 0;
      for (int j = 0; j < n; ++j) s += al[j] * QQ[(size_t)j * n + i];
      rr += al[i] * s;
    }
    double c = sr / sqrt(ss * rr);
    c = 1.0 - fabs(1 - c);
    if (corr) corr[a] = c;
    if (mse) mse[a] = fabs(ss + rr - 2 * sr) / ss;
  }
  if (shape_cov) memcpy(shape_cov, scov.data(), sizeof(double) * scov.size());
  return AAR_OK;
}

extern "C" int aar_sym_eig_extremes(aar_ctx* ctx, int m, int nmat, const double* mats, double* ext) {
  if (!ctx || !mats || !ext || m < 1 || nmat < 1) return AAR_ERR_BAD_ARG;
  AAR_HIP(hipSetDevice(aar_ctx_device(ctx)));
  hipStream_t st = aar_ctx_stream(ctx);
  Dev d;
  double *mats_d, *ext_d;
  int rc;
  const size_t nm = (size_t)m * m * nmat;
  if ((rc = d.alloc(&mats_d, nm)) || (rc = d.alloc(&ext_d, 2 * (size_t)nmat))) return rc;
  AAR_HIP(hipMemcpyAsync(mats_d, mats, sizeof(double) * nm, hipMemcpyHostToDevice, st));
  if ((rc = modal_extremes_launch(mats_d, m, nmat, ext_d, st))) return rc;
  AAR_HIP(hipMemcpyAsync(ext, ext_d, sizeof(double) * 2 * nmat, hipMemcpyDeviceToHost, st));
  AAR_HIP(hipStreamSynchronize(st));
  return AAR_OK;
}
