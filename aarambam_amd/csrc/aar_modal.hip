// Modal decomposition on the device: the mode-bispectrum covariance
// QQ[n][m] = sum_t w_t Q_n(t) Q_m(t) and the shape projections
// QS[s][n] = sum_t S[s][t] w_t Q_n(t) of Aarambam/basis/tetrapyd.c:111-281,
// with Q_n(t) the symmetrised product of three 1-D mode functions at the
// tetrapyd point t (k1 = k[i1], k2 = k[i2], k3 = k[i3]):
//   Q_n = (q1(k1) q2(k2) q3(k3) + q1(k1) q2(k3) q3(k2) + q1(k2) q2(k1) q3(k3)
//          + q1(k2) q2(k3) q3(k1) + q1(k3) q2(k1) q3(k2) + q1(k3) q2(k2) q3(k1)) / 6.
//
// QQ is a dense contraction over the tetrapyd points, so it runs on the FP64
// matrix cores (v_mfma_f64_16x16x4f64).  Q is never stored: for each chunk of
// points a block gathers the 1-D modes at the chunk's k1, k2, k3 into LDS and
// builds the 64 x 32 slices of w Q (rows n) and Q (rows m) from them, then
// feeds them to the MFMA.
// The point range is split over blocks (split-K); the per-split partial
// matrices are summed in a fixed order, so results are deterministic.
#include <hip/hip_runtime.h>

#include "aar_common.h"
#include "aar_kernels.h"

typedef double v4d __attribute__((ext_vector_type(4)));

namespace {

constexpr int kTile = 128;  // output tile (modes n) x (modes m)
constexpr int kChunk = 16;  // tetrapyd points per LDS slice
constexpr int kQuad = kTile / 2;    // each wavefront owns a kQuad x kQuad quadrant
constexpr int kSub = kQuad / 16;    // 16 x 16 MFMA tiles per quadrant side
constexpr int kThreads = 256;

__device__ __forceinline__ double sym_q(const double* ev, int nk, int p1, int p2, int p3, int i1, int i2,
                                        int i3) {
  const double m11 = ev[p1 * nk + i1], m12 = ev[p1 * nk + i2], m13 = ev[p1 * nk + i3];
  const double m21 = ev[p2 * nk + i1], m22 = ev[p2 * nk + i2], m23 = ev[p2 * nk + i3];
  const double m31 = ev[p3 * nk + i1], m32 = ev[p3 * nk + i2], m33 = ev[p3 * nk + i3];
  // times 1/6 instead of the reference's division by 6: at most one ulp apart
  return (m11 * m22 * m33 + m11 * m23 * m32 + m12 * m21 * m33 + m12 * m23 * m31 + m13 * m21 * m32 +
          m13 * m22 * m31) *
         (1.0 / 6.0);
}

// One kTile x kTile tile (tn <= tm, upper triangle of tiles) over the point
// range [t0, t1) of split `sp`: partial[sp][n][m] for the tile.  Four
// wavefronts, each a quadrant of kSub x kSub MFMA tiles of 16 x 16.  Per chunk of points
// the block first gathers every 1-D mode at the chunk's k1, k2, k3 into LDS
// (E[j][p][c]), so building a slice value is nine LDS reads and the products.
__global__ void __launch_bounds__(kThreads)
    k_mode_cov(const double* __restrict__ w, const int* __restrict__ ti1, const int* __restrict__ ti2,
               const int* __restrict__ ti3, int npts, const int* __restrict__ mp1, const int* __restrict__ mp2,
               const int* __restrict__ mp3, int nmodes, const double* __restrict__ evals, int pmax, int nk,
               int ntiles, int nsplit, double* __restrict__ partial) {
  extern __shared__ double smem[];
  double* As = smem;                          // [kTile][kChunk + 1]: w Q for rows n
  double* Bs = As + kTile * (kChunk + 1);     // [kTile][kChunk + 1]: Q for rows m
  double* Ev = Bs + kTile * (kChunk + 1);     // [3][pmax][kChunk]: mode p at k_j of point c
  double* wc = Ev + 3 * pmax * kChunk;        // [kChunk]
  int* pn = (int*)(wc + kChunk);              // [2][3][kTile] mode indices of the tile's rows
  // upper-triangle tile pair from the linear block index
  int tn = 0, rem = blockIdx.x;
  while (rem >= ntiles - tn) {
    rem -= ntiles - tn;
    ++tn;
  }
  const int tm = tn + rem;
  const int sp = blockIdx.y;
  const long long per = ((long long)npts + nsplit - 1) / nsplit;
  const int t0 = (int)(sp * per), t1 = (int)((sp + 1) * per < npts ? (sp + 1) * per : npts);
  for (int q = threadIdx.x; q < 2 * kTile; q += kThreads) {
    const int side = q / kTile, r = q % kTile;
    const int n = (side ? tm : tn) * kTile + r;
    const bool ok = n < nmodes;
    pn[(side * 3 + 0) * kTile + r] = ok ? mp1[n] : -1;
    pn[(side * 3 + 1) * kTile + r] = ok ? mp2[n] : 0;
    pn[(side * 3 + 2) * kTile + r] = ok ? mp3[n] : 0;
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int qr = (wv >> 1) * kQuad, qc = (wv & 1) * kQuad;  // this wavefront's quadrant
  v4d acc[kSub][kSub];
#pragma unroll
  for (int a = 0; a < kSub; ++a)
#pragma unroll
    for (int b = 0; b < kSub; ++b) acc[a][b] = (v4d){0, 0, 0, 0};
  for (int tc = t0; tc < t1; tc += kChunk) {
    // the modes at the chunk's points (zero past t1)
    for (int q = threadIdx.x; q < 3 * pmax * kChunk; q += kThreads) {
      const int j = q / (pmax * kChunk), p = (q / kChunk) % pmax, c = q % kChunk;
      const int t = tc + c;
      double v = 0;
      if (t < t1) v = evals[p * nk + (j == 0 ? ti1[t] : (j == 1 ? ti2[t] : ti3[t]))];
      Ev[q] = v;
    }
    for (int c = threadIdx.x; c < kChunk; c += kThreads) wc[c] = tc + c < t1 ? w[tc + c] : 0.0;
    __syncthreads();
    // the slices: 2 x kTile x kChunk values
    for (int q = threadIdx.x; q < 2 * kTile * kChunk; q += kThreads) {
      const int side = q / (kTile * kChunk), rr = (q / kChunk) % kTile, c = q % kChunk;
      const int p1 = pn[(side * 3 + 0) * kTile + rr];
      double v = 0;
      if (p1 >= 0) {
        const int p2 = pn[(side * 3 + 1) * kTile + rr], p3 = pn[(side * 3 + 2) * kTile + rr];
        const double* E1 = Ev + c;
        const double* E2 = Ev + pmax * kChunk + c;
        const double* E3 = Ev + 2 * pmax * kChunk + c;
        const double m11 = E1[p1 * kChunk], m12 = E2[p1 * kChunk], m13 = E3[p1 * kChunk];
        const double m21 = E1[p2 * kChunk], m22 = E2[p2 * kChunk], m23 = E3[p2 * kChunk];
        const double m31 = E1[p3 * kChunk], m32 = E2[p3 * kChunk], m33 = E3[p3 * kChunk];
        v = (m11 * m22 * m33 + m11 * m23 * m32 + m12 * m21 * m33 + m12 * m23 * m31 + m13 * m21 * m32 +
             m13 * m22 * m31) *
            (1.0 / 6.0);
        if (side == 0) v *= wc[c];
      }
      (side ? Bs : As)[rr * (kChunk + 1) + c] = v;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < kChunk; ks += 4) {
      const int kk = ks + lane / 16;
      double av[kSub], bv[kSub];
#pragma unroll
      for (int a = 0; a < kSub; ++a) av[a] = As[(qr + a * 16 + lane % 16) * (kChunk + 1) + kk];
#pragma unroll
      for (int b = 0; b < kSub; ++b) bv[b] = Bs[(qc + b * 16 + lane % 16) * (kChunk + 1) + kk];
#pragma unroll
      for (int a = 0; a < kSub; ++a)
#pragma unroll
        for (int b = 0; b < kSub; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
    }
    __syncthreads();
  }
  // D[i][j] of a 16 x 16 MFMA tile: lane l, element r -> i = l / 16 + 4 r, j = l % 16
  double* out = partial + (long long)sp * nmodes * nmodes;
#pragma unroll
  for (int a = 0; a < kSub; ++a)
#pragma unroll
    for (int b = 0; b < kSub; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = tn * kTile + qr + a * 16 + lane / 16 + 4 * r;
        const int m = tm * kTile + qc + b * 16 + lane % 16;
        if (n < nmodes && m < nmodes) out[(long long)n * nmodes + m] = acc[a][b][r];
      }
}

// cov[n][m] = sum over splits of partial[sp][min(n,m) tile rows...]: the tiles
// hold n-tile <= m-tile; the lower triangle mirrors the upper.
__global__ void k_mode_cov_reduce(const double* __restrict__ partial, int nsplit, int nmodes,
                                  double* __restrict__ cov) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)nmodes * nmodes) return;
  int n = (int)(e / nmodes), m = (int)(e % nmodes);
  if (n / kTile > m / kTile) {
    const int tmp = n;
    n = m;
    m = tmp;
  }
  double s = 0;
  for (int sp = 0; sp < nsplit; ++sp) s += partial[(long long)sp * nmodes * nmodes + (long long)n * nmodes + m];
  cov[e] = s;
}

// QS partials: one block per (mode, point split); the threads stride over
// the split's points and each w Q_n(t) is used for every shape s.
__global__ void __launch_bounds__(kThreads)
    k_mode_qs(const double* __restrict__ S, int nshapes, const double* __restrict__ w, const int* __restrict__ ti1,
              const int* __restrict__ ti2, const int* __restrict__ ti3, int npts, const int* __restrict__ mp1,
              const int* __restrict__ mp2, const int* __restrict__ mp3, int nmodes, const double* __restrict__ evals,
              int pmax, int nk, int nsplit, double* __restrict__ partial) {
  extern __shared__ double smem[];
  double* ev = smem;
  double* red = ev + pmax * nk;  // [kThreads]
  for (int q = threadIdx.x; q < pmax * nk; q += kThreads) ev[q] = evals[q];
  __syncthreads();
  const int n = blockIdx.x;  // one mode per block
  const int sp = blockIdx.y;
  const long long per = ((long long)npts + nsplit - 1) / nsplit;
  const int t0 = (int)(sp * per), t1 = (int)((sp + 1) * per < npts ? (sp + 1) * per : npts);
  const int p1 = mp1[n], p2 = mp2[n], p3 = mp3[n];
  for (int s = 0; s < nshapes; ++s) {
    double acc = 0;
    for (int t = t0 + threadIdx.x; t < t1; t += kThreads)
      acc += S[(long long)s * npts + t] * (w[t] * sym_q(ev, nk, p1, p2, p3, ti1[t], ti2[t], ti3[t]));
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int h = kThreads / 2; h > 0; h >>= 1) {
      if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[((long long)sp * nshapes + s) * nmodes + n] = red[0];
    __syncthreads();
  }
}

__global__ void k_mode_qs_reduce(const double* __restrict__ partial, int nsplit, int nshapes, int nmodes,
                                 double* __restrict__ qs) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nshapes * nmodes) return;
  double s = 0;
  for (int sp = 0; sp < nsplit; ++sp) s += partial[(long long)sp * nshapes * nmodes + e];
  qs[e] = s;
}

// Residual of the expansion at every tetrapyd point, R[t] = S[t] - sum_j c_j Q_{n_j}(t)
// (c_j = alpha_j / norm_j, modes n_j given by their index triples), and per block
// the partial sum of w_t R[t]^2.
__global__ void __launch_bounds__(kThreads)
    k_mode_residual(const double* __restrict__ S, const double* __restrict__ w, const int* __restrict__ ti1,
                    const int* __restrict__ ti2, const int* __restrict__ ti3, int npts,
                    const int* __restrict__ sel, int nsel, const double* __restrict__ coef,
                    const int* __restrict__ mp1, const int* __restrict__ mp2, const int* __restrict__ mp3,
                    const double* __restrict__ evals, int pmax, int nk, double* __restrict__ R,
                    double* __restrict__ partial) {
  extern __shared__ double smem[];
  double* ev = smem;
  double* cf = ev + pmax * nk;  // [nsel]
  int* sp = reinterpret_cast<int*>(cf + nsel);  // [3][nsel]
  double* red = reinterpret_cast<double*>(sp + 3 * ((nsel + 1) & ~1));  // [kThreads]
  for (int q = threadIdx.x; q < pmax * nk; q += kThreads) ev[q] = evals[q];
  for (int j = threadIdx.x; j < nsel; j += kThreads) {
    const int n = sel[j];
    cf[j] = coef[j];
    sp[j] = mp1[n];
    sp[nsel + j] = mp2[n];
    sp[2 * nsel + j] = mp3[n];
  }
  __syncthreads();
  const int t = blockIdx.x * kThreads + threadIdx.x;
  double acc = 0;
  if (t < npts) {
    const int i1 = ti1[t], i2 = ti2[t], i3 = ti3[t];
    double sum = 0;
    for (int j = 0; j < nsel; ++j) sum += sym_q(ev, nk, sp[j], sp[nsel + j], sp[2 * nsel + j], i1, i2, i3) * cf[j];
    const double r = S[t] - sum;
    R[t] = r;
    acc = r * w[t] * r;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int h = kThreads / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// The bordered Gram matrices of the candidate modes: matrix b (column-major,
// m = nsel + 1 rows) holds qq[idx][idx] for idx = the selected modes followed
// by candidate cand[b].
__global__ void k_mode_bordered(const double* __restrict__ qq, int nmodes, const int* __restrict__ sel, int nsel,
                                const int* __restrict__ cand, double* __restrict__ out) {
  const int m = nsel + 1;
  const int b = blockIdx.y;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)m * m) return;
  const int r = (int)(e % m), c = (int)(e / m);
  const int ir = r < nsel ? sel[r] : cand[b], ic = c < nsel ? sel[c] : cand[b];
  out[(long long)b * m * m + e] = qq[(long long)ir * nmodes + ic];
}

// Sum (max when kMax) of one value per thread over the block; every thread gets
// the result.  The tree has a fixed shape, so the result is deterministic.
template <bool kMax>
__device__ double block_all(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) {
    const double u = __shfl_xor(v, o);
    v = kMax ? fmax(v, u) : v + u;
  }
  const int wave = threadIdx.x >> 6;
  __syncthreads();  // red may still be read by the previous call
  if ((threadIdx.x & 63) == 0) red[wave] = v;
  __syncthreads();
  v = red[0];
  for (int q = 1; q < kThreads / 64; ++q) v = kMax ? fmax(v, red[q]) : v + red[q];
  return v;
}

// Number of eigenvalues of the symmetric tridiagonal matrix (d, off-diagonal
// squares e2) below x: the signs of the LDL^T pivots of T - x I (Sturm count by
// Kahan's recurrence, a pivot smaller than pivmin taken as -pivmin as LAPACK's
// dstebz does).
__device__ int sturm_count(const double* d, const double* e2, int m, double x, double pivmin) {
  double q = d[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  int c = q < 0;
  for (int i = 1; i < m; ++i) {
    q = (d[i] - x) - e2[i - 1] / q;
    if (fabs(q) < pivmin) q = -pivmin;
    c += q < 0;
  }
  return c;
}

// The k-th smallest eigenvalue (1-based) of the tridiagonal by multisection:
// the wavefront's 64 lanes count at 64 evenly spaced points of [lo, hi] and
// keep the sub-interval where the count reaches k, until the interval is
// within two ulps of its end points.  Invariant: count(lo) < k <= count(hi).
__device__ double kth_eigenvalue(const double* d, const double* e2, int m, int k, double lo, double hi,
                                 double pivmin) {
  const int lane = threadIdx.x & 63;
  for (int it = 0; it < 64; ++it) {
    const double tol = 2 * __DBL_EPSILON__ * fmax(fabs(lo), fabs(hi)) + 2 * pivmin;
    if (hi - lo <= tol) break;
    const double step = (hi - lo) / 65;
    const double x = lo + (lane + 1) * step;
    const unsigned long long hit = __ballot(sturm_count(d, e2, m, x, pivmin) >= k);
    if (hit == 0) {
      lo = __shfl(x, 63);
    } else {
      const int j = __builtin_ctzll(hit);
      const double xl = __shfl(x, j > 0 ? j - 1 : 0);
      hi = __shfl(x, j);
      if (j > 0) lo = xl;
    }
  }
  return 0.5 * (lo + hi);
}

// Smallest and largest eigenvalue magnitude of each symmetric m x m matrix of
// mats (column-major, both triangles, overwritten): ext[2b] = min |l|,
// ext[2b + 1] = max |l|.  For a symmetric matrix these are its extreme singular
// values, whose ratio is np.linalg.cond's 2-norm condition number.
// One block per matrix: Householder reduction to tridiagonal form (Golub and
// Van Loan, Matrix Computations, 4th ed., Algorithm 8.3.1, with house() of
// Algorithm 5.1.1), then bisection for eigenvalues by index, one index per
// wavefront: the smallest, the largest, and the two either side of zero.
// kLds: the reduction works on a copy of the matrix in LDS (m^2 doubles), else
// in place in global memory.
template <bool kLds>
__global__ void __launch_bounds__(kThreads) k_mode_extremes(double* __restrict__ mats, int m, double* __restrict__ ext) {
  extern __shared__ double smem[];
  double* Ag = mats + (size_t)blockIdx.x * m * m;
  double* A = kLds ? smem : Ag;
  double* v = smem + (kLds ? (size_t)m * m : 0);  // [m] Householder vector
  double* wv = v + m;        // [m] p, then w
  double* d = wv + m;        // [m] diagonal
  double* e2 = d + m;        // [m] squared off-diagonal
  double* part = e2 + m;     // [kThreads] matvec partial sums
  double* red = part + kThreads;  // [kThreads / 64]
  const int t = threadIdx.x;
  if (kLds) {
    for (int e = t; e < m * m; e += kThreads) A[e] = Ag[e];
    __syncthreads();
  }
  for (int k = 0; k + 2 < m; ++k) {
    const int n2 = m - k - 1, o = k + 1;  // trailing block: rows and columns o .. m - 1
    const double* x = A + (size_t)k * m + o;
    double s = 0;
    for (int i = 1 + t; i < n2; i += kThreads) s += x[i] * x[i];
    const double sigma = block_all<false>(s, red);
    const double alpha = x[0];
    if (t == 0) d[k] = A[(size_t)k * m + k];
    if (sigma == 0) {  // column already reduced: H = I
      if (t == 0) e2[k] = alpha * alpha;
      continue;
    }
    const double mu = sqrt(alpha * alpha + sigma);
    const double beta = alpha <= 0 ? alpha - mu : -sigma / (alpha + mu);
    const double tau = 2 * beta * beta / (sigma + beta * beta);
    if (t == 0) e2[k] = mu * mu;
    for (int i = t; i < n2; i += kThreads) v[i] = i == 0 ? 1.0 : x[i] / beta;
    __syncthreads();
    // p = tau A22 v: rows over the threads, the columns split into nsl slices
    const int nsl = n2 <= kThreads ? kThreads / n2 : 1;
    if (nsl > 1) {
      const int i = t % n2, sl = t / n2;
      if (sl < nsl) {
        double acc = 0;
        for (int j = sl; j < n2; j += nsl) acc += A[(size_t)(o + j) * m + o + i] * v[j];
        part[sl * n2 + i] = acc;
      }
      __syncthreads();
      if (t < n2) {
        double acc = 0;
        for (int q = 0; q < nsl; ++q) acc += part[q * n2 + t];
        wv[t] = tau * acc;
      }
    } else {
      for (int i = t; i < n2; i += kThreads) {
        double acc = 0;
        for (int j = 0; j < n2; ++j) acc += A[(size_t)(o + j) * m + o + i] * v[j];
        wv[i] = tau * acc;
      }
    }
    __syncthreads();
    double pv = 0;
    for (int i = t; i < n2; i += kThreads) pv += wv[i] * v[i];
    const double c = 0.5 * tau * block_all<false>(pv, red);
    for (int i = t; i < n2; i += kThreads) wv[i] -= c * v[i];  // w = p - (tau p.v / 2) v
    __syncthreads();
    // A22 -= v w^T + w v^T
    if (nsl > 1) {
      const int i = t % n2, sl = t / n2;
      if (sl < nsl)
        for (int j = sl; j < n2; j += nsl) A[(size_t)(o + j) * m + o + i] -= v[i] * wv[j] + wv[i] * v[j];
    } else {
      for (int i = t; i < n2; i += kThreads)
        for (int j = 0; j < n2; ++j) A[(size_t)(o + j) * m + o + i] -= v[i] * wv[j] + wv[i] * v[j];
    }
    __syncthreads();
  }
  if (t == 0) {
    if (m >= 2) {
      d[m - 2] = A[(size_t)(m - 2) * m + m - 2];
      const double e = A[(size_t)(m - 2) * m + m - 1];
      e2[m - 2] = e * e;
    }
    d[m - 1] = A[(size_t)(m - 1) * m + m - 1];
  }
  __syncthreads();
  // Gershgorin interval and the pivot floor (LAPACK dstebz's pivmin and widening)
  double lo = INFINITY, hi = -INFINITY, emax = 0;
  for (int i = t; i < m; i += kThreads) {
    const double r = (i > 0 ? sqrt(e2[i - 1]) : 0.0) + (i + 1 < m ? sqrt(e2[i]) : 0.0);
    lo = fmin(lo, d[i] - r);
    hi = fmax(hi, d[i] + r);
    if (i + 1 < m) emax = fmax(emax, e2[i]);
  }
  lo = -block_all<true>(-lo, red);
  hi = block_all<true>(hi, red);
  const double pivmin = __DBL_MIN__ * fmax(1.0, block_all<true>(emax, red));
  const double tnorm = fmax(fabs(lo), fabs(hi));
  lo -= 2.1 * tnorm * __DBL_EPSILON__ * m + 4.2 * pivmin;
  hi += 2.1 * tnorm * __DBL_EPSILON__ * m + 4.2 * pivmin;
  const int nneg = sturm_count(d, e2, m, 0.0, pivmin);
  const int wave = t >> 6;
  const int idx[4] = {1, m, nneg > 1 ? nneg : 1, nneg < m ? nneg + 1 : m};
  const double l = m == 1 ? d[0] : kth_eigenvalue(d, e2, m, idx[wave], lo, hi, pivmin);
  if ((t & 63) == 0) part[wave] = fabs(l);  // part is idle since the reduction's last barrier
  __syncthreads();
  if (t == 0) {
    ext[2 * blockIdx.x] = fmin(part[2], part[3]);
    ext[2 * blockIdx.x + 1] = fmax(part[0], part[1]);
  }
}

}  // namespace

int modal_residual_launch(const double* S, const double* w, const int* i1, const int* i2, const int* i3, int npts,
                          const int* sel, int nsel, const double* coef, const int* p1, const int* p2, const int* p3,
                          const double* evals, int pmax, int nk, double* R, double* partial, hipStream_t st) {
  const size_t lds = sizeof(double) * ((size_t)pmax * nk + nsel + kThreads) + sizeof(int) * 3 * ((nsel + 1) & ~1);
  if (lds > 64 * 1024) return AAR_ERR_UNSUPPORTED;
  const int nb = (npts + kThreads - 1) / kThreads;
  hipLaunchKernelGGL(k_mode_residual, dim3(nb), dim3(kThreads), lds, st, S, w, i1, i2, i3, npts, sel, nsel, coef,
                     p1, p2, p3, evals, pmax, nk, R, partial);
  return hipGetLastError() == hipSuccess ? AAR_OK : AAR_ERR_HIP;
}

int modal_bordered_launch(const double* qq, int nmodes, const int* sel, int nsel, const int* cand, int ncand,
                          double* out, hipStream_t st) {
  const long long mm = (long long)(nsel + 1) * (nsel + 1);
  if (ncand < 1 || ncand > 65535) return AAR_ERR_UNSUPPORTED;
  hipLaunchKernelGGL(k_mode_bordered, dim3((unsigned)((mm + 255) / 256), ncand), dim3(256), 0, st, qq, nmodes, sel,
                     nsel, cand, out);
  return hipGetLastError() == hipSuccess ? AAR_OK : AAR_ERR_HIP;
}

int modal_cov_launch(const double* w, const int* i1, const int* i2, const int* i3, int npts, const int* p1,
                     const int* p2, const int* p3, int nmodes, const double* evals, int pmax, int nk, int nsplit,
                     double* partial, double* cov, hipStream_t st) {
  const int ntiles = (nmodes + kTile - 1) / kTile;
  const int npairs = ntiles * (ntiles + 1) / 2;
  const size_t lds = sizeof(double) * (2 * kTile * (kChunk + 1) + 3 * (size_t)pmax * kChunk + kChunk) +
                     sizeof(int) * 6 * kTile;
  if (lds > 64 * 1024) return AAR_ERR_UNSUPPORTED;
  hipLaunchKernelGGL(k_mode_cov, dim3(npairs, nsplit), dim3(kThreads), lds, st, w, i1, i2, i3, npts, p1, p2, p3,
                     nmodes, evals, pmax, nk, ntiles, nsplit, partial);
  if (hipGetLastError() != hipSuccess) return AAR_ERR_HIP;
  const long long ne = (long long)nmodes * nmodes;
  hipLaunchKernelGGL(k_mode_cov_reduce, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, st, partial, nsplit,
                     nmodes, cov);
  return hipGetLastError() == hipSuccess ? AAR_OK : AAR_ERR_HIP;
}

int modal_qs_launch(const double* S, int nshapes, const double* w, const int* i1, const int* i2, const int* i3,
                    int npts, const int* p1, const int* p2, const int* p3, int nmodes, const double* evals, int pmax,
                    int nk, int nsplit, double* partial, double* qs, hipStream_t st) {
  const size_t lds = sizeof(double) * ((size_t)pmax * nk + kThreads);
  if (lds > 64 * 1024) return AAR_ERR_UNSUPPORTED;
  hipLaunchKernelGGL(k_mode_qs, dim3(nmodes, nsplit), dim3(kThreads), lds, st, S, nshapes, w, i1, i2, i3, npts, p1,
                     p2, p3, nmodes, evals, pmax, nk, nsplit, partial);
  if (hipGetLastError() != hipSuccess) return AAR_ERR_HIP;
  hipLaunchKernelGGL(k_mode_qs_reduce, dim3((unsigned)((nshapes * nmodes + 255) / 256)), dim3(256), 0, st, partial,
                     nsplit, nshapes, nmodes, qs);
  return hipGetLastError() == hipSuccess ? AAR_OK : AAR_ERR_HIP;
}

int modal_extremes_launch(double* mats, int m, int nmat, double* ext, hipStream_t st) {
  const size_t lds = sizeof(double) * (4 * (size_t)m + kThreads + kThreads / 64);
  if (m < 1 || nmat < 1 || lds > 64 * 1024) return AAR_ERR_UNSUPPORTED;
  // the matrix itself in LDS when it fits the device's per-block LDS (160 KiB on
  // gfx950: m <= 140)
  int dev = 0, lds_max = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
    return AAR_ERR_HIP;
  const size_t lds_in = lds + sizeof(double) * (size_t)m * m;
  if (lds_in <= (size_t)lds_max &&
      hipFuncSetAttribute((const void*)k_mode_extremes<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)lds_in) == hipSuccess) {
    hipLaunchKernelGGL(k_mode_extremes<true>, dim3(nmat), dim3(kThreads), lds_in, st, mats, m, ext);
  } else {
    hipLaunchKernelGGL(k_mode_extremes<false>, dim3(nmat), dim3(kThreads), lds, st, mats, m, ext);
  }
  return hipGetLastError() == hipSuccess ? AAR_OK : AAR_ERR_HIP;
}
