// Device-side types and helpers of the MI355X NG-2LPT library: complex arithmetic, the grid
// geometry and its row order, the mesh-size rule and the twiddle-table layout.  Free of host
// headers, so the kernels compiled at run time (hiprtc, aar_jit.cpp) include it as well.
#pragma once
#ifndef __HIPCC_RTC__  // hiprtc provides the HIP device types and intrinsics itself
#include <hip/hip_runtime.h>
#endif

#define AAR_PI 3.14159265358979323846  // Aarambam/LPT/Basis/allvars.h:4

template <typename T> struct Cx;
template <> struct Cx<float> { using t = float2; };
template <> struct Cx<double> { using t = double2; };
template <typename T> using cx_t = typename Cx<T>::t;

template <typename C> __host__ __device__ __forceinline__ C cmake(decltype(C::x) a, decltype(C::x) b) {
  C r;
  r.x = a;
  r.y = b;
  return r;
}
template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return cmake<C>(a.x + b.x, a.y + b.y); }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return cmake<C>(a.x - b.x, a.y - b.y); }
// Complex product with two fused multiply-adds (4 instructions instead of 6;
// the build turns off automatic contraction, so this one is explicit).
template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
  return cmake<C>(fma(a.x, b.x, -(a.y * b.y)), fma(a.x, b.y, a.y * b.x));
}
template <typename C> __device__ __forceinline__ C cconj(C a) { return cmake<C>(a.x, -a.y); }
template <typename C> __device__ __forceinline__ C cscale(C a, decltype(C::x) s) { return cmake<C>(a.x * s, a.y * s); }
// multiply by DIR*i (DIR = +1 or -1)
template <int DIR, typename C> __device__ __forceinline__ C cmul_i(C a) {
  return DIR > 0 ? cmake<C>(-a.y, a.x) : cmake<C>(a.y, -a.x);
}

// Grid geometry of one field in HBM.  Grid rows (x, y) are stored x-blocked:
// row r = ((x / 8) * N + y) * 8 + x % 8, so eight consecutive x share each
// stretch of memory and a column along x crosses N/8 memory regions instead
// of N (fewer pages and DRAM rows per x-pass tile).  A spectrum holds the
// kz < N/2 bins as main[r][kz] (N*N*H complex, H = N/2) followed by the
// kz = N/2 bins as nyq[r] (N*N complex).  The same buffer holds a real field
// as real[r][z] (N^3 reals) in the bytes of the main block.
struct Geom {
  int N, H;
  int nlog, hlog;  // log2 N, log2 H when N is a power of two
  int pow2;        // N is a power of two (else the mixed-radix FFT and divisions decode rows)
  long long nq;  // offset (complex elements) of the Nyquist plane = N*N*H
  long long fs;  // field stride (complex elements) = N*N*(H+1)
};

// x-planes per block of the row order
#ifndef AAR_XB_LOG
#define AAR_XB_LOG 3
#endif
constexpr int kXBLog = AAR_XB_LOG, kXB = 1 << kXBLog;
__host__ __device__ __forceinline__ long long row_of(int N, int x, int y) {
  return ((long long)(x >> kXBLog) * N + y) * kXB + (x & (kXB - 1));
}
__host__ __device__ __forceinline__ void row_xy(int N, long long r, int& x, int& y) {
  const long long t = r >> kXBLog;
  y = (int)(t % N);
  x = (int)(t / N) * kXB + (int)(r & (kXB - 1));
}
// same, for power-of-two N given as log2 (no 64-bit division)
__host__ __device__ __forceinline__ void row_xy_log(int nlog, unsigned r, int& x, int& y) {
  const unsigned t = r >> kXBLog;
  y = (int)(t & ((1u << nlog) - 1));
  x = (int)((t >> nlog) << kXBLog) + (int)(r & (kXB - 1));
}

// Mesh sizes: powers of two from 16 to 1024, and, through the mixed-radix FFT,
// the other multiples of 16 up to 1024 whose prime factors are 2, 3, 5 and 7
// (48, 80, 96, 112, 144, 160, 192, ..., 768, ..., 1008).  N % 16 == 0 keeps
// every x-block of 8 planes whole and N/2 a multiple of the Gaussian kernel's
// 8-mode chunks.
__host__ __device__ constexpr bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }
__host__ __device__ constexpr bool smooth7(int n) {
  for (int p = 2; p <= 7; ++p)
    while (n % p == 0) n /= p;
  return n == 1;
}
__host__ __device__ constexpr bool mesh_ok(int n) { return n >= 16 && n <= 1024 && n % 16 == 0 && smooth7(n); }

// Twiddle tables: exp(-2 pi i m / L) at tw[L + m] for the powers of two L <= 2048
// (4096 entries), then one table of N entries per mesh size N that is not a
// power of two, in increasing N, at tw_offset(N).
constexpr int kTwPow2 = 4096;
__host__ __device__ constexpr int tw_offset(int n) {
  if (is_pow2(n)) return n;
  int off = kTwPow2;
  for (int m = 16; m < n; m += 16)
    if (mesh_ok(m) && !is_pow2(m)) off += m;
  return off;
}
constexpr int kTwTotal = tw_offset(1024 + 16) > kTwPow2 ? tw_offset(1024 + 16) : kTwPow2;

static inline Geom make_geom(int N) {
  Geom g;
  g.N = N;
  g.H = N / 2;
  g.pow2 = is_pow2(N);
  g.nlog = 0;
  while ((1 << g.nlog) < N) ++g.nlog;
  g.hlog = g.nlog - 1;
  g.nq = (long long)N * N * (N / 2);
  g.fs = (long long)N * N * (N / 2 + 1);
  return g;
}
