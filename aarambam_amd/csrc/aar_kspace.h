// k-space helpers shared by the pipeline and statistics kernels: element ->
// (x, y, kz) decoding, the Nsample cut, |k|^2 from the per-axis tables, the
// Legendre recurrence and the radial weights of the modal basis.
#pragma once
#include "aar_kernels.h"

__device__ __forceinline__ void decode_k(const Geom& g, long long e, int& x, int& y, int& kz) {
  // N*N rows fit 32 bits for N <= 1024 (2^20 rows), and so does a whole field
  unsigned r;
  if (g.pow2) {
    if (e < g.nq) {
      kz = (int)(e & (g.H - 1));
      r = (unsigned)(e >> g.hlog);
    } else {
      kz = g.H;
      r = (unsigned)(e - g.nq);
    }
    row_xy_log(g.nlog, r, x, y);
  } else {
    const unsigned ue = (unsigned)e, uh = (unsigned)g.H, un = (unsigned)g.N;
    if (e < g.nq) {
      r = ue / uh;
      kz = (int)(ue - r * uh);
    } else {
      kz = g.H;
      r = (unsigned)(e - g.nq);
    }
    const unsigned t = r >> kXBLog;
    const unsigned xb = t / un;
    y = (int)(t - xb * un);
    x = (int)(xb << kXBLog) + (int)(r & (kXB - 1));
  }
}

// Runs body(x, y, kz) for element e of a block of consecutive elements.  When the block lies in
// one row of the kz < N/2 block (N/2 a multiple of the block size) the row, hence x and y, comes
// from the block index: the body then sees wave-uniform x and y, and its table reads indexed by
// them become scalar loads.  Otherwise each element is decoded on its own.
template <class F> __device__ __forceinline__ void with_k(const Geom& g, long long e, F&& body) {
  const long long e0 = (long long)blockIdx.x * blockDim.x;
  if (g.H % (int)blockDim.x == 0 && e0 < g.nq) {
    int x, y, z0;
    decode_k(g, e0, x, y, z0);
    body(x, y, (int)(e - e0) + z0);
  } else {
    int x, y, z;
    decode_k(g, e, x, y, z);
    body(x, y, z);
  }
}

// main.c:452-462: SphereMode 1 keeps |k| L/2pi <= Nsample/2, else the cube.
__device__ __forceinline__ bool outside_cut(const KSpace& ks, int x, int y, int z, double kmag1) {
  if (ks.sphere == 1) return kmag1 * ks.box / (2 * AAR_PI) > ks.nsh;
  return (ks.outc[x] | ks.outc[y] | ks.outc[z]) != 0;
}

__device__ __forceinline__ double kmag2_of(const KSpace& ks, int x, int y, int z) {
  const double a = ks.kv[x], b = ks.kv[y], c = ks.kv[z];
  return a * a + b * b + c * c;
}

// Oscillation chain of the ResBasis Gaussian potential (ResBasis/main.c:445-455),
// applied to Anorm k^ns in this order.  Its switch on WhichSpectrum has no
// breaks: 0 and 1 apply the factor in k and the factor in ln k, 2 the one in
// ln k only (res bits 1|2, 1|2 and 2; 0 for Basis).
__device__ __forceinline__ double res_modulate(const KSpace& ks, double v, double kmag) {
  if (ks.res & 1) v *= (1.0 + ks.res_A * cos(ks.res_w * kmag + ks.res_phi));
  if (ks.res & 2) v *= (1.0 + ks.res_A * cos(ks.res_w * log(kmag) + ks.res_phi));
  return v;
}

// Upward Legendre recurrence, GSL gsl_sf_legendre_Pl (legendre_poly.c).
// k^s of a mode, s = (4 - n_s)/3, from |k|^2 (FP32) or |k| (FP64).  The mode fields and the
// combine weights both take it from here, so on the FP32 path their Legendre arguments agree:
// 2^(s/2 log2 |k|^2) with the hardware exp2 / log2; the FP64 path keeps pow(|k|, s).
template <typename RT> __device__ __forceinline__ RT kpow_s(double kmag2, double kmag1, double s) {
  if constexpr (sizeof(RT) == 4) return exp2f(0.5f * (float)s * log2f((float)kmag2));
  else return pow(kmag1, s);
}

// RT is the precision of the radial functions: FP64 on the FP64 path, FP32 on
// the FP32 path (|x| <= 1 and orders <= 31 keep the recurrence stable).
template <typename RT> struct LegRecT {
  RT x, pm2, pm1;
  int l;
  __device__ void init(RT xx) {
    x = xx;
    pm2 = 1;
    pm1 = xx;
    l = 1;
  }
  __device__ RT at(int L) {
    if (L == 0) return 1;
    while (l < L) {
      ++l;
      RT p = (x * (RT)(2 * l - 1) * pm1 - (RT)(l - 1) * pm2) / (RT)l;
      pm2 = pm1;
      pm1 = p;
    }
    return pm1;
  }
};
using LegRec = LegRecT<double>;

// Radial weight g_n(k) of the modal basis (main.c:835-862):
// n <= 3: k^{s(n-3)};  n >= 4: (P_{n-1}(xbar) - X0[n-1]) / k^2.
template <typename RT>
__device__ __forceinline__ RT weight_g(int n, RT ks, RT kmag2, LegRecT<RT>& leg, const double* x0) {
  if (n <= 3) {
    RT r = 1;
    for (int q = n; q < 3; ++q) r *= ks;
    return (RT)1 / r;
  }
  return (leg.at(n - 1) - (RT)x0[n - 1]) / kmag2;
}

