// Binned power spectrum and bispectrum of a job's potentials, for checking
// Phi_NG against its input template.
//
// Every statistic is a sum over real space of products of band-filtered
// fields: with F_b(x) = C2R[f(k) 1_b(|k|)] (unnormalised, FFTW convention),
//   sum_x F_b(x) G_b(x)            = N^3 sum_{k in b} f(k) g(k)
//   sum_x F_b1(x) G_b2(x) H_b3(x)  = N^3 sum_{k1+k2+k3=0, ki in bi} f g h,
// so one kernel builds the filtered spectra, the FP64 C2R of the pipeline
// transforms them, and one kernel forms the weighted product sums.
#include <math.h>

#include "aar_kernels.h"
#include "aar_kspace.h"

// kind 0: (a - b) * scale;  1: 1;  2: P(k) q_n(k) (n < 0: P(k));  3: g_n(k);  4: |k|
// with P(k) = pamp k^ns / k^4 (times the ResBasis oscillations), the variance
// of phi_G(k).  Bins are
// |k| in [edges[b], edges[b+1]); the origin, the Nyquist planes and the modes
// outside the Nsample cut belong to no bin.
template <typename TS>
__global__ void __launch_bounds__(256)
    k_stat_bins(const cx_t<TS>* __restrict__ a, const cx_t<TS>* __restrict__ b, double scale, int kind, int n,
                double pamp, double ns, double2* __restrict__ out, int nb, const double* __restrict__ edges, Geom g,
                KSpace ks) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= g.fs) return;
  int x, y, z;
  decode_k(g, e, x, y, z);
  const double k2 = kmag2_of(ks, x, y, z), k1 = sqrt(k2);
  int bin = -1;
  if (!(x == 0 && y == 0 && z == 0) && x != g.H && y != g.H && z != g.H && !outside_cut(ks, x, y, z, k1))
    for (int q = 0; q < nb; ++q)
      if (k1 >= edges[q] && k1 < edges[q + 1]) {
        bin = q;
        break;
      }
  double2 v = make_double2(0, 0);
  if (bin >= 0) {
    if (kind == 0) {
      double re = a[e].x, im = a[e].y;
      if (b) {
        re -= (double)b[e].x;
        im -= (double)b[e].y;
      }
      v = make_double2(re * scale, im * scale);
    } else if (kind == 1) {
      v = make_double2(1, 0);
    } else if (kind == 4) {
      v = make_double2(k1, 0);
    } else {
      const double kss = pow(k1, ks.s);
      LegRec leg;
      leg.init(-1 + 2 * (kss - ks.kmin_s) / (ks.kmax_s - ks.kmin_s));
      double r;
      if (kind == 2) {
        double q = 1;
        if (n >= 4) q = (leg.at(n - 1) - ks.x0[n - 1]) * (kss * kss * kss / k2);
        else
          for (int t = 0; t < n; ++t) q *= kss;
        r = res_modulate(ks, pamp * pow(k1, ns), k1) / (k2 * k2) * q;
      } else {
        r = weight_g<double>(n, kss, k2, leg, ks.x0);
      }
      v = make_double2(r, 0);
    }
  }
  for (int q = 0; q < nb; ++q) out[(long long)q * g.fs + e] = q == bin ? v : make_double2(0, 0);
}

template <typename TS>
int stat_bin_fields(const cx_t<TS>* a, const cx_t<TS>* b, double scale, int kind, int n, double pamp, double ns,
                    double2* out, int nb, const double* edges, const Geom& g, const KSpace& ks, hipStream_t st) {
  hipLaunchKernelGGL((k_stat_bins<TS>), dim3((unsigned)((g.fs + 255) / 256)), dim3(256), 0, st, a, b, scale, kind,
                     n, pamp, ns, out, nb, edges, g, ks);
  return hipGetLastError() == hipSuccess ? AAR_OK : AAR_ERR_HIP;
}
template int stat_bin_fields<float>(const float2*, const float2*, double, int, int, double, double, double2*, int,
                                    const double*, const Geom&, const KSpace&, hipStream_t);
template int stat_bin_fields<double>(const double2*, const double2*, double, int, int, double, double, double2*,
                                     int, const double*, const Geom&, const KSpace&, hipStream_t);

// partial[o * nchunks + c] = sum over terms t of output o of
//   w_t sum_{x in chunk c} A_t(x) B_t(x) (C_t(x) or 1 when fc < 0)
// Real field f is ((const double*)fields)[f * 2 fs + x] for x < N^3.
__global__ void __launch_bounds__(256)
    k_stat_sums(const double* __restrict__ fields, long long fstride, long long n3, int nchunks,
                const int* __restrict__ ofs, const StatTerm* __restrict__ terms, double* __restrict__ partial) {
  const int o = blockIdx.y, c = blockIdx.x;
  const long long per = (n3 + nchunks - 1) / nchunks;
  const long long lo = c * per, hi = lo + per < n3 ? lo + per : n3;
  double acc = 0;
  for (int t = ofs[o]; t < ofs[o + 1]; ++t) {
    const StatTerm tm = terms[t];
    const double* A = fields + tm.a * fstride;
    const double* B = fields + tm.b * fstride;
    const double* Cf = tm.c >= 0 ? fields + tm.c * fstride : nullptr;
    double s = 0;
    for (long long x = lo + threadIdx.x; x < hi; x += blockDim.x) s += A[x] * B[x] * (Cf ? Cf[x] : 1.0);
    acc += tm.w * s;
  }
  __shared__ double red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[(long long)o * nchunks + c] = red[0];
}

int stat_sums(const double2* fields, const Geom& g, const int* ofs, const StatTerm* terms, int nout, int nchunks,
              double* partial, hipStream_t st) {
  if (nout < 1 || nout > 65535 || nchunks < 1) return AAR_ERR_BAD_ARG;
  hipLaunchKernelGGL(k_stat_sums, dim3(nchunks, nout), dim3(256), 0, st, (const double*)fields, 2 * g.fs,
                     (long long)g.N * g.N * g.N, nchunks, ofs, terms, partial);
  return hipGetLastError() == hipSuccess ? AAR_OK : AAR_ERR_HIP;
}
