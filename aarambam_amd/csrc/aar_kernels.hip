// HIP kernels of the NG-2LPT initial-condition pipeline for gfx950.
//
// All k-space kernels evaluate the reference's expressions in the same
// operation order (compiled with -ffp-contract=off), reading the per-axis
// k components and cut flags from host-built tables, so that cut decisions
// and radial arguments agree bit for bit with Aarambam/LPT/Basis/main.c.
#include <math.h>

#include <atomic>

#include "aar_fft.h"
#include "aar_kernels.h"
#include "aar_zng.h"
#include "aar_kspace.h"
#include "aar_ranlxd.h"

// ------------------------------------------------------------------ twiddles
template <typename T> void make_twiddles(cx_t<T>* tw) {
  const long double tp = 6.283185307179586476925286766559005768L;
  tw[0] = cmake<cx_t<T>>(0, 0);
  tw[1] = cmake<cx_t<T>>(0, 0);
  auto fill = [&](int L, int at) {
    for (int m = 0; m < L; ++m) {
      long double a = tp * (long double)m / (long double)L;
      tw[at + m] = cmake<cx_t<T>>((T)cosl(a), (T)(-sinl(a)));
    }
  };
  for (int L = 2; L <= 2048; L *= 2) fill(L, L);
  for (int n = 16; n <= 1024; n += 16)
    if (mesh_ok(n) && !is_pow2(n)) fill(n, tw_offset(n));
}
template void make_twiddles<float>(float2*);
template void make_twiddles<double>(double2*);

// ------------------------------------------------------------------ launch shapes
template <typename T> struct Shape {
  static constexpr int CMAX = sizeof(T) == 4 ? 16 : 8;  // 128-byte row segments
};
__host__ __device__ constexpr int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
// Threads per tile: 16 elements per thread where possible, at least 2.
__host__ __device__ constexpr int tile_threads(int CL) {
  return clampi(CL / 16, CL / 2 < 64 ? CL / 2 : 64, 1024);
}
// Tile shapes of the x / y passes, picked per precision, mesh size and axis
// by the autotuner.  All keep 16 elements per thread, hence the same radix
// plan and bit-identical results; they differ in columns per tile:
//   V = 0: 128-byte column segments;
//   V = 1: 64-byte segments (half the LDS and threads per tile, so more
//          tiles per CU at L = 1024; the XCD-aware block order keeps both
//          halves of each 128-byte line on one XCD's L2);
//   V = 2: 256-byte segments (fewer, larger tiles; L <= 512);
//   V = 3: 128-byte segments, four tiles in turn per workgroup, the next
//          tile's loads in flight during this tile's transform (FP32,
//          L >= 512: at L = 1024 one 16-wave tile fills a CU's registers,
//          so without it the loads and the transform never overlap).
template <typename T, int L, int V = 0> struct XYCfg {
  static constexpr int TPW = V == 3 ? 4 : 1;  // tiles per workgroup
  static constexpr int CW = V == 1 ? Shape<T>::CMAX / 2 : (V == 2 ? Shape<T>::CMAX * 2 : Shape<T>::CMAX);
  static constexpr int C = CW < L / 2 ? CW : L / 2;
  static constexpr bool VALID =
      C * L / 16 <= 1024 && (long)L * (C + 1) * sizeof(T) + (long)L * sizeof(cx_t<T>) <= 152 * 1024 &&
      (V != 3 || (sizeof(T) == 4 && L >= 512));
  static constexpr int NT = tile_threads(C * L);
  // minimum wavefronts per SIMD the register allocation must allow: 64-byte
  // tiles at L = 1024 fit three per CU by LDS when they stay within 80 VGPRs
  static constexpr int WPE = (sizeof(T) == 4 && L == 1024 && V == 1) ? 6 : 1;
};

// Variant table: [precision 0 FP32 / 1 FP64][log2 N][axis]
// Tuning tables, read by every launch and written by the autotuner or a pin from any thread.
// Only powers of two have entries; other mesh sizes run the mixed-radix passes (0 everywhere).
// Axis 2 is the y pass of the multi-output pairs (xy_split), whose loads also apply the output
// factors; -1 there follows axis 1.
static std::atomic<int> g_xy_variant[2][11][3] = {{{0, 0, -1}, {0, 0, -1}, {0, 0, -1}, {0, 0, -1}, {0, 0, -1},
                                                   {0, 0, -1}, {0, 0, -1}, {0, 0, -1}, {0, 0, -1}, {0, 0, -1},
                                                   {0, 0, -1}},
                                                  {{0, 0, -1}, {0, 0, -1}, {0, 0, -1}, {0, 0, -1}, {0, 0, -1},
                                                   {0, 0, -1}, {0, 0, -1}, {0, 0, -1}, {0, 0, -1}, {0, 0, -1},
                                                   {0, 0, -1}}};
int aar_xy_variant(int fp64, int n, int axis) {
  if (!is_pow2(n)) return 0;
  int l = 0;
  while ((1 << l) < n) ++l;
  if (l >= 11 || axis < 0 || axis > 2) return 0;
  const int v = g_xy_variant[fp64 ? 1 : 0][l][axis].load(std::memory_order_relaxed);
  return v >= 0 ? v : g_xy_variant[fp64 ? 1 : 0][l][1].load(std::memory_order_relaxed);
}
void aar_set_xy_variant(int fp64, int n, int axis, int v) {
  if (!is_pow2(n)) return;
  int l = 0;
  while ((1 << l) < n) ++l;
  if (l < 11 && axis >= 0 && axis < 3 && v >= (axis == 2 ? -1 : 0) && v < kXYVariantCount)
    g_xy_variant[fp64 ? 1 : 0][l][axis].store(v, std::memory_order_relaxed);
}

// Radix plans of the x / y passes: [precision][log2 N] (aar_fft.h plan_radix)
static std::atomic<int> g_xy_plan[2][11] = {};
int aar_xy_plan(int fp64, int n) {
  if (!is_pow2(n)) return 0;
  int l = 0;
  while ((1 << l) < n) ++l;
  return l < 11 ? g_xy_plan[fp64 ? 1 : 0][l].load(std::memory_order_relaxed) : 0;
}
void aar_set_xy_plan(int fp64, int n, int plan) {
  if (!is_pow2(n)) return;
  int l = 0;
  while ((1 << l) < n) ++l;
  if (l < 11 && plan >= 0 && plan < kXYPlanCount) g_xy_plan[fp64 ? 1 : 0][l].store(plan, std::memory_order_relaxed);
}

// Column groups of the x / y pass pairs: [precision][log2 N], 0 = whole passes over all fields
static std::atomic<int> g_xy_group[2][11] = {};
int aar_xy_group(int fp64, int n) {
  if (!is_pow2(n)) return 0;
  int l = 0;
  while ((1 << l) < n) ++l;
  return l < 11 ? g_xy_group[fp64 ? 1 : 0][l].load(std::memory_order_relaxed) : 0;
}
void aar_set_xy_group(int fp64, int n, int cols) {
  if (!is_pow2(n)) return;
  int l = 0;
  while ((1 << l) < n) ++l;
  if (l < 11 && cols >= 0) g_xy_group[fp64 ? 1 : 0][l].store(cols, std::memory_order_relaxed);
}
// Column groups of the multi-output x / y pass pairs (xy_split), as g_xy_group; -1: the pair's
static std::atomic<int> g_xy_split_group[2][11] = {{-1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1},
                                                   {-1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1}};
int aar_xy_split_group(int fp64, int n) {
  if (!is_pow2(n)) return 0;
  int l = 0;
  while ((1 << l) < n) ++l;
  if (l >= 11) return 0;
  const int v = g_xy_split_group[fp64 ? 1 : 0][l].load(std::memory_order_relaxed);
  return v >= 0 ? v : aar_xy_group(fp64, n);
}
void aar_set_xy_split_group(int fp64, int n, int cols) {
  if (!is_pow2(n)) return;
  int l = 0;
  while ((1 << l) < n) ++l;
  if (l < 11 && cols >= -1) g_xy_split_group[fp64 ? 1 : 0][l].store(cols, std::memory_order_relaxed);
}
int aar_xy_pair_dispatches(int nf, int H, int group_cols) {
  return group_cols > 0 ? 2 * nf * ((H + group_cols - 1) / group_cols) : 2;
}


// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X; give
// each XCD a contiguous range of tiles so neighbouring tiles share its L2.
__device__ __forceinline__ int xcd_tile(int b, int nblk) {
  if (nblk % 8) return b;
  return (b % 8) * (nblk / 8) + b / 8;
}
template <typename T, int H> struct ZCfg {
  // FP32 H = 512: 8 elements per thread (the radix-8 plan of 16 elements per
  // thread, with one row per wavefront); otherwise 16 per thread
  static constexpr bool E8 = sizeof(T) == 4 && H == 512;
  static constexpr int C = E8 ? 8 : Shape<T>::CMAX;
  static constexpr int NT = E8 ? C * H / 8 : tile_threads(C * H);
  // every row's exchanges inside one wavefront: no workgroup barriers
  static constexpr bool WL = wave_local(H, C * H / NT) && NT % 64 == 0;
};

// ------------------------------------------------------------------ x / y passes
// One tile = C adjacent columns of one field.  Tiles [0, nmain) cover the
// kz < N/2 block, the rest the Nyquist plane.
// MO = 1: the y pass of the xo.n outputs of a multi-output split (XSplit),
// workgroup w (in the XCD-aware order) doing pair w % xo.n of tiles w / xo.n,
// read from its input's slot times its factor of (y, kz): the pairs of one tile
// run side by side on one XCD, so an input read by several outputs comes from
// its L2 after the first.
// MO = 2: the x pass of the xm.M mode fields of phi_G (XModes), workgroup w doing
// mode w % M of tiles w / M, phi_G read times the mode's filter q_n(|k|).
template <typename T, int L, int AXIS, int DIR, int V, int PL, int MO = 0>
__global__ void __launch_bounds__((XYCfg<T, L, V>::NT)) __attribute__((amdgpu_waves_per_eu(XYCfg<T, L, V>::WPE)))
    k_xy_pass(cx_t<T>* __restrict__ base, Geom g, const cx_t<T>* __restrict__ twg, int nmain, int gper, int k0,
              XOut xo, XModes xm) {
  using C2 = cx_t<T>;
  constexpr int C = XYCfg<T, L, V>::C, NT = XYCfg<T, L, V>::NT;
  static_assert(MO != 1 || (AXIS == 1 && DIR > 0), "multi-output: inverse y passes");
  static_assert(MO != 2 || (AXIS == 0 && DIR > 0), "mode filters: inverse x passes");
  __shared__ T tile_r[L * (C + 1)];  // real-valued exchange tile (one component at a time)
  __shared__ C2 tw[L];
  __shared__ T fyt[MO == 1 ? L : 1];  // MO 1: the output's factor of ky, per y
  load_twiddles(tw, twg + L, L, NT);
  // MO 1: this block's output, its input and its code, picked with constant indices (the index
  // is uniform; a runtime index into the argument struct would go through scratch)
  int src = blockIdx.y, dst = blockIdx.y;
  unsigned code = 0;
  const int w = xcd_tile(blockIdx.x, gridDim.x);  // this workgroup's (tile block, output)
  const int nouts = MO == 1 ? xo.n : (MO == 2 ? xm.M : 1);
  const int wt = MO ? w / nouts : w;
  int nmode = 0;  // MO 2: the mode number of this block's output
  if constexpr (MO == 2) {
    dst = w % nouts;
    nmode = xm.modes[dst];
  }
  if constexpr (MO == 1) {
    const int q = w % xo.n;
#pragma unroll
    for (int i = 0; i < 9; ++i)
      if (i == q) {
        src = xo.src[i];
        dst = xo.dst[i];
        code = xo.code[i];
      }
    for (int y = threadIdx.x; y < L; y += NT) {
      double fy = 1;
      if (code & kXsKy) fy *= xo.kv[y];
      if (code & kXsKy2) fy *= xo.kv[y] * xo.kv[y];
      if (code & kXsCic) fy *= xo.icic[y];
      fyt[y] = (T)fy;
    }
  }
  const C2* fi = MO == 2 ? (const C2*)xm.phig : base + (long long)src * g.fs;
  C2* fo = base + (long long)dst * g.fs;
  constexpr int TPW = XYCfg<T, L, V>::TPW;
  const int N = g.N, H = g.H;
  // MO 1: the factor of kz (with the sign) of this thread's column of tile t (y pass: the
  // column's kz, or kz = N/2 on the Nyquist plane)
  auto factor_z = [&](int t) -> T {
    const int zi = t < nmain ? k0 + (t % gper) * C + (int)(threadIdx.x % C) : H;
    double fz = (code & kXsNeg) ? -1.0 : 1.0;
    if (code & kXsKz) fz *= xo.kv[zi];
    if (code & kXsKz2) fz *= xo.kv[zi] * xo.kv[zi];
    if (code & kXsCic) fz *= xo.icic[zi];
    return (T)fz;
  };
  // The offset of element (pos, col) of tile t is (pos / B) s1 + (pos % B) s2 + ct, B = kXB,
  // with tile-uniform strides s1, s2 and a column term ct that the thread computes once per
  // tile (each thread of the split-exchange tile keeps one column, col = threadIdx.x % C).
  // Offsets fit 32 bits: a field holds N^2 (N/2 + 1) < 2^31 complex values for N <= 1024.
  struct Addr {
    unsigned s1, s2, ct;
  };
  auto tile_addr = [&](int t) {
    Addr A;
    const int col = threadIdx.x % C;
    if (t < nmain) {
      const int a = t / gper;             // fixed y (x pass) or fixed x (y pass)
      const int c0 = k0 + (t % gper) * C;  // first kz of the tile
      A.s1 = AXIS == 0 ? (unsigned)(kXB * N * H) : (unsigned)(kXB * kXB * H);
      A.s2 = AXIS == 0 ? (unsigned)H : (unsigned)(kXB * H);
      // row_of(N, pos, a) * H + c0 + col, or row_of(N, a, pos) = row_of(N, a, 0) + kXB pos
      A.ct = AXIS == 0 ? (unsigned)(kXB * a * H + c0 + col) : (unsigned)(row_of(N, a, 0) * H + c0 + col);
    } else {
      const int c0 = (t - nmain) * C;  // first y (x pass) or first x (y pass) of a Nyquist tile
      A.s1 = AXIS == 0 ? (unsigned)(kXB * N) : (unsigned)(kXB * kXB);
      A.s2 = AXIS == 0 ? 1u : (unsigned)kXB;
      // nq + row_of(N, pos, c0 + col), or nq + row_of(N, c0 + col, pos)
      A.ct = AXIS == 0 ? (unsigned)(g.nq + (long long)kXB * (c0 + col)) : (unsigned)(g.nq + row_of(N, c0 + col, 0));
    }
    return A;
  };
  auto off = [](const Addr& A, int pos) -> unsigned {
    return (unsigned)(pos >> kXBLog) * A.s1 + (unsigned)(pos & (kXB - 1)) * A.s2 + A.ct;
  };
  // MO 2: the (y, kz) grid indices of this thread's column of tile t and their parts of |k|^2 and
  // of the cut (x pass: the tile's y, the column's kz, or kz = N/2 on the Nyquist plane)
  struct ColK {
    int yz;          // y + kz
    double kyy, kzz;  // kv[y]^2, kv[kz]^2
    int oyz;         // cube cut of y or kz
  };
  auto col_k = [&](int t) {
    ColK c;
    const int col = threadIdx.x % C;
    int yi, zi;
    if (t < nmain) {
      yi = t / gper;
      zi = k0 + (t % gper) * C + col;
    } else {
      yi = (t - nmain) * C + col;
      zi = H;
    }
    c.yz = yi + zi;
    const double b = xm.ks.kv[yi], cz = xm.ks.kv[zi];
    c.kyy = b * b;
    c.kzz = cz * cz;
    c.oyz = xm.ks.outc[yi] | xm.ks.outc[zi];
    return c;
  };
  // MO 2: phi_G at (x = pos, y, kz) times q_n (k_modes_spec's arithmetic: the same operations in
  // the same order, so the mode fields are those of k_modes_spec followed by an x pass)
  auto mode_filter = [&](const ColK& c, int pos, C2 v) -> C2 {
    using RT = T;
    const KSpace& ks = xm.ks;
    const double a = ks.kv[pos];
    const double kmag2 = a * a + c.kyy + c.kzz;
    const double kmag1 = (sizeof(T) == 8 || ks.sphere == 1) ? sqrt(kmag2) : 0.0;
    const bool cut = ks.sphere == 1 ? kmag1 * ks.box / (2 * AAR_PI) > ks.nsh : (ks.outc[pos] | c.oyz) != 0;
    if (pos + c.yz == 0 || cut) return cmake<C2>(0, 0);
    const RT k2 = (RT)kmag2;
    const RT kss = kpow_s<RT>(kmag2, kmag1, ks.s);
    RT q;
    if (nmode <= 3) {
      q = 1;
      for (int u = 0; u < nmode; ++u) q *= kss;
    } else {
      const RT xb = -1 + 2 * (kss - (RT)ks.kmin_s) / (RT)(ks.kmax_s - ks.kmin_s);
      const RT k3s = kss * kss * kss;  // k^{4-ns}
      LegRecT<RT> leg;
      leg.init(xb);
      q = (leg.at(nmode - 1) - (RT)ks.x0[nmode - 1]) * (k3s / k2);
    }
    return cmake<C2>(q * v.x, q * v.y);
  };
  // the tile's input element at pos (MO 1: times fz fy(pos), and i where the code says so)
  auto load = [&](const Addr& A, T fz, int pos) -> C2 {
    const C2 v = fi[off(A, pos)];
    if constexpr (MO == 1) {
      const T s = fz * fyt[pos];
      return (code & kXsI) ? cmake<C2>(-s * v.y, s * v.x) : cmake<C2>(s * v.x, s * v.y);
    } else {
      return v;
    }
  };
  if constexpr (TPW == 1) {
    const int t = wt;
    const Addr A = tile_addr(t);
    const T fz = MO == 1 ? factor_z(t) : (T)1;
    __syncthreads();  // twiddles (and the factor table)
    if constexpr (MO == 2) {
      const ColK ck = col_k(t);
      fft_tile_split<T, L, C, NT, DIR, PL>(
          tile_r, tw, [&](int pos, int) { return mode_filter(ck, pos, fi[off(A, pos)]); },
          [&](int pos, int, C2 v) { fo[off(A, pos)] = v; });
    } else {
      fft_tile_split<T, L, C, NT, DIR, PL>(
          tile_r, tw, [&](int pos, int) { return load(A, fz, pos); }, [&](int pos, int, C2 v) { fo[off(A, pos)] = v; });
    }
  } else {
    // tiles t0 .. t0 + TPW - 1 (the last workgroup may hold fewer); the condition of every
    // branch and loop below is uniform over the workgroup, so every barrier is reached
    constexpr int E = C * L / NT;
    const int ntiles = nmain + N / C;
    const int t0 = wt * TPW;
    // MO: the next tile's loads stay raw while in flight; its factors (or filters) are applied
    // when it is taken up (factors of the slots in flight would hold E more registers)
    C2 v[E], nv[E];
    Addr A = tile_addr(t0);
    if constexpr (MO == 1) __syncthreads();  // the factor table
    split_load<T, L, C, NT, PL>(nv, [&](int pos, int) { return fi[off(A, pos)]; });
    for (int i = 0; i < TPW && t0 + i < ntiles; ++i) {
#pragma unroll
      for (int e = 0; e < E; ++e) v[e] = nv[e];
      const Addr Ac = A;
      if (i + 1 < TPW && t0 + i + 1 < ntiles) {
        A = tile_addr(t0 + i + 1);
        split_load<T, L, C, NT, PL>(nv, [&](int pos, int) { return fi[off(A, pos)]; });
      }
      if constexpr (MO == 2) {
        const ColK ck = col_k(t0 + i);
#pragma unroll
        for (int e = 0; e < E; ++e) {
          int pos, col;
          split_slot<T, L, C, NT, PL>(e, pos, col);
          v[e] = mode_filter(ck, pos, v[e]);
        }
      }
      if constexpr (MO == 1) {
        const T fz = factor_z(t0 + i);
#pragma unroll
        for (int e = 0; e < E; ++e) {
          int pos, col;
          split_slot<T, L, C, NT, PL>(e, pos, col);
          const T sc = fz * fyt[pos];
          v[e] = (code & kXsI) ? cmake<C2>(-sc * v[e].y, sc * v[e].x) : cmake<C2>(sc * v[e].x, sc * v[e].y);
        }
      }
      split_transform<T, L, C, NT, DIR, PL>(tile_r, tw, v, [&](int pos, int, C2 x) { fo[off(Ac, pos)] = x; });
    }
  }
}

// xo: the y pass of the outputs of a multi-output split (inverse y passes only, nf = 1)
template <typename T, int L, int AXIS, int DIR, int V, int PL>
static int launch_xy_v(cx_t<T>* f, int nf, const Geom& g, const cx_t<T>* tw, hipStream_t st, const XYRange& r,
                       const XOut* xo, const XModes* xm) {
  using Cfg = XYCfg<T, L, V>;
  if constexpr (!Cfg::VALID) {
    return AAR_ERR_UNSUPPORTED;
  } else {
    constexpr int C = Cfg::C, NT = Cfg::NT;
    const int cols = r.cols > 0 ? r.cols : g.H;
    if (cols % C || r.k0 % C || r.k0 < 0 || r.k0 + cols > g.H) return AAR_ERR_UNSUPPORTED;
    const int gper = cols / C;
    const int nmain = g.N * gper;
    const int nnyq = r.nyq ? g.N / C : 0;
    if (nmain + nnyq == 0) return AAR_OK;
    const int nwg = (nmain + nnyq + Cfg::TPW - 1) / Cfg::TPW;
    if (xo) {
      // the factor table takes L more values of LDS
      constexpr bool MO_FITS =
          (long)L * (C + 1) * sizeof(T) + (long)L * sizeof(cx_t<T>) + (long)L * sizeof(T) <= 160 * 1024;
      if (nf != 1 || xo->n < 1 || xo->n > 9) return AAR_ERR_UNSUPPORTED;
      if constexpr (AXIS == 1 && DIR > 0 && MO_FITS) {
        hipLaunchKernelGGL((k_xy_pass<T, L, AXIS, DIR, V, PL, 1>), dim3(nwg * xo->n, 1), dim3(NT), 0, st, f, g,
                           tw, nmain, gper, r.k0, *xo, XModes());
      } else {
        return AAR_ERR_UNSUPPORTED;
      }
    } else if (xm) {
      if (nf != 1 || xm->M < 1 || (long long)nwg * xm->M > 0x7fffffffLL) return AAR_ERR_UNSUPPORTED;
      if constexpr (AXIS == 0 && DIR > 0) {
        hipLaunchKernelGGL((k_xy_pass<T, L, AXIS, DIR, V, PL, 2>), dim3(nwg * xm->M, 1), dim3(NT), 0, st, f, g,
                           tw, nmain, gper, r.k0, XOut(), *xm);
      } else {
        return AAR_ERR_UNSUPPORTED;
      }
    } else {
      hipLaunchKernelGGL((k_xy_pass<T, L, AXIS, DIR, V, PL>), dim3(nwg, nf), dim3(NT), 0, st, f, g, tw, nmain, gper,
                         r.k0, XOut(), XModes());
    }
    return hipGetLastError() == hipSuccess ? AAR_OK : AAR_ERR_HIP;
  }
}

// Tile variants other than 0 are built for the sizes where the choice matters; radix plans other
// than 0 for L = 512 and 1024 (three stages of radix 16, 8, 4 or 2 in different orders).
template <typename T, int L, int AXIS, int DIR, int PL>
static int launch_xy_p(cx_t<T>* f, int nf, const Geom& g, const cx_t<T>* tw, hipStream_t st, int v,
                       const XYRange& r, const XOut* xs, const XModes* xm) {
  if constexpr (L >= 256) {
    switch (v) {
      case 1: return launch_xy_v<T, L, AXIS, DIR, 1, PL>(f, nf, g, tw, st, r, xs, xm);
      case 2: return launch_xy_v<T, L, AXIS, DIR, 2, PL>(f, nf, g, tw, st, r, xs, xm);
      case 3: return launch_xy_v<T, L, AXIS, DIR, 3, PL>(f, nf, g, tw, st, r, xs, xm);
      default: break;
    }
  }
  return launch_xy_v<T, L, AXIS, DIR, 0, PL>(f, nf, g, tw, st, r, xs, xm);
}

template <typename T, int L, int AXIS, int DIR>
static int launch_xy(cx_t<T>* f, int nf, const Geom& g, const cx_t<T>* tw, hipStream_t st, int v, const XYRange& r,
                     const XOut* xs, const XModes* xm) {
  if constexpr (L >= 512) {
    switch (aar_xy_plan(sizeof(T) == 8, L)) {
      case 1: return launch_xy_p<T, L, AXIS, DIR, 1>(f, nf, g, tw, st, v, r, xs, xm);
      case 2: return launch_xy_p<T, L, AXIS, DIR, 2>(f, nf, g, tw, st, v, r, xs, xm);
      default: break;
    }
  }
  return launch_xy_p<T, L, AXIS, DIR, 0>(f, nf, g, tw, st, v, r, xs, xm);
}

// ------------------------------------------------------------------ z passes
// Real transforms along the contiguous axis, C rows per block, through a
// half-length complex FFT.  c2r: Y[k] = (X_k + X*_{H-k}) + i (X_k - X*_{H-k}) e^{+2 pi i k/N},
// z = IFFT_H(Y), real row = (Re z_m, Im z_m) interleaved; the imaginary
// parts of X_0 and X_H are ignored (FFTW's half-complex convention).
// r2c: Z = FFT_H(x_{2m} + i x_{2m+1}), X_k = E_k + e^{-2 pi i k/N} O_k.
template <typename T, int H>
__global__ void __launch_bounds__((ZCfg<T, H>::NT))
    k_z_c2r(cx_t<T>* __restrict__ base, Geom g, const cx_t<T>* __restrict__ twg) {
  using C2 = cx_t<T>;
  constexpr int C = ZCfg<T, H>::C, NT = ZCfg<T, H>::NT;
  constexpr int N = 2 * H;
  using TI = TileIdx<T, H, C, true, C * H / NT>;
  __shared__ C2 tile[TI::SIZE];
  __shared__ C2 tw[N];
  __shared__ C2 xh[C];
  load_twiddles(tw, twg + N, N, NT);
  C2* f = base + (long long)blockIdx.y * g.fs;
  const long long r0 = (long long)blockIdx.x * C;
  C2* rows = f + r0 * H;
  if (threadIdx.x < C) xh[threadIdx.x] = f[g.nq + r0 + threadIdx.x];
  __syncthreads();
  // The first FFT stage forms Y[k] from X[k] and X[H-k], read straight from
  // HBM along each row (the mirrored read hits the same cache lines).
  auto load = [&](int k, int c) {
    const C2* row = rows + (long long)c * H;
    if (k == 0) {
      const T a = row[0].x, b = xh[c].x;
      return cmake<C2>(a + b, a - b);
    }
    const C2 A = row[k], B = row[H - k];
    C2 w = tw[k];
    w.y = -w.y;  // e^{+2 pi i k/N}
    const C2 s = cadd(A, cconj(B)), d = csub(A, cconj(B));
    const C2 dw = cmul(d, w);
    return cmake<C2>(s.x - dw.y, s.y + dw.x);
  };
  auto store = [&](int pos, int c, C2 v) { rows[(long long)c * H + pos] = v; };
  // true: barrier after the first gather, since the transform is in place in
  // HBM (with wavefront-local rows, each output already depends on every load)
  fft_tile<T, H, C, NT, +1, 2, true, true, ZCfg<T, H>::WL>(tile, tw, load, store);
}

template <typename T, int H>
__global__ void __launch_bounds__((ZCfg<T, H>::NT))
    k_z_r2c(cx_t<T>* __restrict__ base, Geom g, const cx_t<T>* __restrict__ twg) {
  using C2 = cx_t<T>;
  constexpr int C = ZCfg<T, H>::C, NT = ZCfg<T, H>::NT;
  constexpr int N = 2 * H;
  using TI = TileIdx<T, H, C, true, C * H / NT>;
  __shared__ C2 tile[TI::SIZE];
  __shared__ C2 tw[N];
  load_twiddles(tw, twg + N, N, NT);
  __syncthreads();
  C2* f = base + (long long)blockIdx.y * g.fs;
  const long long r0 = (long long)blockIdx.x * C;
  C2* rows = f + r0 * H;
  auto load = [&](int pos, int c) { return rows[(long long)c * H + pos]; };
  auto store = [&](int pos, int c, C2 v) { tile[TI::at(pos, c)] = v; };
  constexpr bool WL = ZCfg<T, H>::WL;
  fft_tile<T, H, C, NT, -1, 2, false, true, WL>(tile, tw, load, store);
  tile_sync<WL>();
  r2c_post<T, H, C, NT, WL>(tile, tw, f, r0, g);
}

template <typename T, int H>
static int launch_z(cx_t<T>* f, int nf, const Geom& g, const cx_t<T>* tw, int dir, hipStream_t st) {
  constexpr int C = ZCfg<T, H>::C, NT = ZCfg<T, H>::NT;
  dim3 grid((unsigned)((long long)g.N * g.N / C), nf);
  if (dir > 0)
    hipLaunchKernelGGL((k_z_c2r<T, H>), grid, dim3(NT), 0, st, f, g, tw);
  else
    hipLaunchKernelGGL((k_z_r2c<T, H>), grid, dim3(NT), 0, st, f, g, tw);
  return hipGetLastError() == hipSuccess ? AAR_OK : AAR_ERR_HIP;
}

// ------------------------------------------------------------------ mixed-radix passes
// Mesh sizes that are not powers of two: the same passes over the same layout, with the
// mixed-radix Stockham transform (gfft, aar_fft.h) on C interleaved vectors in LDS.  Dynamic
// LDS: the N-point twiddle table, then two buffers of C (L + 1) ... vectors of L elements.
template <typename T> static int gen_cols(int L, int N) {
  for (int C : {8, 4, 2})
    if ((2LL * L * (C + 1) + N) * (long long)sizeof(cx_t<T>) <= 150 * 1024) return C;
  return 2;
}
template <typename T> static size_t gen_lds(int L, int N, int C) {
  return (2 * (size_t)L * (C + 1) + N) * sizeof(cx_t<T>);
}
constexpr int kGenThreads = 512;

// x or y pass: tile t of C adjacent kz columns (t < nmain) or C adjacent rows of the Nyquist
// plane, field blockIdx.y; twg: the N-point table (tw_offset(N))
template <typename T, int C, int AXIS, int DIR>
__global__ void __launch_bounds__(kGenThreads)
    k_xy_gen(cx_t<T>* __restrict__ base, Geom g, const cx_t<T>* __restrict__ twg, GPlan pl, int nmain) {
  using C2 = cx_t<T>;
  constexpr int CP = C + 1;
  extern __shared__ char smem[];
  const int N = g.N, H = g.H;
  C2* tw = reinterpret_cast<C2*>(smem);
  C2* a = tw + N;
  C2* b = a + (size_t)N * CP;
  for (int m = threadIdx.x; m < N; m += kGenThreads) tw[m] = twg[m];
  C2* f = base + (long long)blockIdx.y * g.fs;
  const int t = blockIdx.x;
  const bool main_tile = t < nmain;
  const int per = H / C;
  const int q = main_tile ? t / per : 0, kz0 = main_tile ? (t % per) * C : 0;  // q: fixed y (x) or x (y)
  const int rc0 = main_tile ? 0 : (t - nmain) * C;  // first fixed y (x pass) or x (y pass) of a Nyquist tile
  // offset of element pos of column c of this tile
  auto off = [&](int pos, int c) -> long long {
    if (main_tile) return (AXIS == 0 ? row_of(N, pos, q) : row_of(N, q, pos)) * H + kz0 + c;
    return g.nq + (AXIS == 0 ? row_of(N, pos, rc0 + c) : row_of(N, rc0 + c, pos));
  };
  for (int q = threadIdx.x; q < C * N; q += kGenThreads) {
    const int c = q % C, pos = q / C;
    a[pos * CP + c] = f[off(pos, c)];
  }
  __syncthreads();
  const C2* r = gfft<T, DIR, C>(a, b, tw, 1, pl, kGenThreads);
  for (int q = threadIdx.x; q < C * N; q += kGenThreads) {
    const int c = q % C, pos = q / C;
    f[off(pos, c)] = r[pos * CP + c];
  }
}

template <typename T, int C, int AXIS, int DIR>
static int launch_xy_gen_c(cx_t<T>* f, int nf, const Geom& g, const cx_t<T>* tw, hipStream_t st) {
  const GPlan pl = make_gplan(g.N);
  if (!pl.ns || g.H % C || g.N % C) return AAR_ERR_UNSUPPORTED;
  const size_t lds = gen_lds<T>(g.N, g.N, C);
  static const hipError_t attr = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_xy_gen<T, C, AXIS, DIR>),
                                                     hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024);
  if (attr != hipSuccess) return AAR_ERR_HIP;
  const int nmain = g.N * (g.H / C);
  dim3 grid(nmain + g.N / C, nf);
  hipLaunchKernelGGL((k_xy_gen<T, C, AXIS, DIR>), grid, dim3(kGenThreads), lds, st, f, g, tw + tw_offset(g.N), pl,
                     nmain);
  return hipGetLastError() == hipSuccess ? AAR_OK : AAR_ERR_HIP;
}

template <typename T, int AXIS, int DIR>
static int launch_xy_gen(cx_t<T>* f, int nf, const Geom& g, const cx_t<T>* tw, hipStream_t st) {
  switch (gen_cols<T>(g.N, g.N)) {
    case 8: return launch_xy_gen_c<T, 8, AXIS, DIR>(f, nf, g, tw, st);
    case 4: return launch_xy_gen_c<T, 4, AXIS, DIR>(f, nf, g, tw, st);
    default: return launch_xy_gen_c<T, 2, AXIS, DIR>(f, nf, g, tw, st);
  }
}

// z pass over C rows (blockIdx.x) of field blockIdx.y through an H-point transform: the
// half-complex <-> real conversions of k_z_c2r / k_z_r2c around gfft (H-point twiddles are
// every second entry of the N-point table)
template <typename T, int C, int DIR>
__global__ void __launch_bounds__(kGenThreads)
    k_z_gen(cx_t<T>* __restrict__ base, Geom g, const cx_t<T>* __restrict__ twg, GPlan pl) {
  using C2 = cx_t<T>;
  constexpr int CP = C + 1;
  extern __shared__ char smem[];
  const int N = g.N, H = g.H;
  C2* tw = reinterpret_cast<C2*>(smem);
  C2* a = tw + N;
  C2* b = a + (size_t)H * CP;
  for (int m = threadIdx.x; m < N; m += kGenThreads) tw[m] = twg[m];
  __syncthreads();
  C2* f = base + (long long)blockIdx.y * g.fs;
  const long long r0 = (long long)blockIdx.x * C;
  C2* rows = f + r0 * H;
  if constexpr (DIR > 0) {
    // Y[k] = (X_k + X*_{H-k}) + i (X_k - X*_{H-k}) e^{+2 pi i k/N}; X_H from the Nyquist plane
    for (int c = 0; c < C; ++c)
    for (int k = threadIdx.x; k < H; k += kGenThreads) {
      const C2* row = rows + (long long)c * H;
      C2 y;
      if (k == 0) {
        const T x0 = row[0].x, xh = f[g.nq + r0 + c].x;
        y = cmake<C2>(x0 + xh, x0 - xh);
      } else {
        const C2 A = row[k], B = row[H - k];
        C2 w = tw[k];
        w.y = -w.y;
        const C2 sm = cadd(A, cconj(B)), d = csub(A, cconj(B));
        const C2 dw = cmul(d, w);
        y = cmake<C2>(sm.x - dw.y, sm.y + dw.x);
      }
      a[k * CP + c] = y;
    }
    __syncthreads();
    const C2* r = gfft<T, +1, C>(a, b, tw, 2, pl, kGenThreads);
    for (int c = 0; c < C; ++c)
      for (int m = threadIdx.x; m < H; m += kGenThreads) rows[(long long)c * H + m] = r[m * CP + c];
  } else {
    for (int c = 0; c < C; ++c)
      for (int m = threadIdx.x; m < H; m += kGenThreads) a[m * CP + c] = rows[(long long)c * H + m];
    __syncthreads();
    const C2* r = gfft<T, -1, C>(a, b, tw, 2, pl, kGenThreads);
    // X_k = E_k + e^{-2 pi i k/N} O_k from Z_k and Z_{H-k}; X_0 and the Nyquist entry from Z_0
    for (int c = 0; c < C; ++c)
    for (int k = threadIdx.x; k < H; k += kGenThreads) {
      C2* out = rows + (long long)c * H;
      if (k == 0) {
        const C2 z = r[c];
        out[0] = cmake<C2>(z.x + z.y, (T)0);
        f[g.nq + r0 + c] = cmake<C2>(z.x - z.y, (T)0);
      } else {
        const C2 A = r[k * CP + c], B = r[(H - k) * CP + c];
        const C2 Ev = cscale(cadd(A, cconj(B)), (T)0.5);
        const C2 dd = csub(A, cconj(B));
        const C2 O = cmake<C2>(dd.y * (T)0.5, -dd.x * (T)0.5);
        out[k] = cadd(Ev, cmul(O, tw[k]));
      }
    }
  }
}

template <typename T, int C, int DIR>
static int launch_z_gen_c(cx_t<T>* f, int nf, const Geom& g, const cx_t<T>* tw, hipStream_t st) {
  const GPlan pl = make_gplan(g.H);
  if (!pl.ns || ((long long)g.N * g.N) % C) return AAR_ERR_UNSUPPORTED;
  const size_t lds = gen_lds<T>(g.H, g.N, C);
  static const hipError_t attr = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_z_gen<T, C, DIR>),
                                                     hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024);
  if (attr != hipSuccess) return AAR_ERR_HIP;
  dim3 grid((unsigned)((long long)g.N * g.N / C), nf);
  hipLaunchKernelGGL((k_z_gen<T, C, DIR>), grid, dim3(kGenThreads), lds, st, f, g, tw + tw_offset(g.N), pl);
  return hipGetLastError() == hipSuccess ? AAR_OK : AAR_ERR_HIP;
}

template <typename T>
static int launch_z_gen(cx_t<T>* f, int nf, const Geom& g, const cx_t<T>* tw, int dir, hipStream_t st) {
  const int C = gen_cols<T>(g.H, g.N);
  if (dir > 0) {
    if (C == 8) return launch_z_gen_c<T, 8, +1>(f, nf, g, tw, st);
    if (C == 4) return launch_z_gen_c<T, 4, +1>(f, nf, g, tw, st);
    return launch_z_gen_c<T, 2, +1>(f, nf, g, tw, st);
  }
  if (C == 8) return launch_z_gen_c<T, 8, -1>(f, nf, g, tw, st);
  if (C == 4) return launch_z_gen_c<T, 4, -1>(f, nf, g, tw, st);
  return launch_z_gen_c<T, 2, -1>(f, nf, g, tw, st);
}

// ------------------------------------------------------------------ fused z pass of the NG sum
// (the body, z_ng_body, is in aar_zng.h; the kernels compiled at run time share it)
// Products of the NG sum from the factored records (the greedy vertex cover of each output
// group, aar_job.cpp ng_factor): one 64-byte scalar load per block of four records, issued
// before the header is examined; every read of a mode slot is a run of indexed register reads
// with one uniform index.
template <typename T> struct RecordProducts {
  const int* gs;
  const NgRec4* recs;
  template <int E, class Get>
  __device__ __forceinline__ void operator()(int kk, Get& get, cx_t<T>* hv) const {
    using C2 = cx_t<T>;
    typedef int blk_t __attribute__((ext_vector_type(16)));
    auto coef = [](const blk_t& k, int q) -> T {
      if constexpr (sizeof(T) == 4) {
        return __int_as_float(k[4 * q + 1]);
      } else {
        return __hiloint2double(k[4 * q + 3], k[4 * q + 2]);
      }
    };
    for (int r = gs[kk]; r < gs[kk + 1];) {
      const blk_t* R = reinterpret_cast<const blk_t*>(recs + (r >> 2));
      blk_t blk = R[0];
      const int v = blk[0], n = blk[1];
      C2 u[E], b[E];
#pragma unroll
      for (int e = 0; e < E; ++e) u[e] = cmake<C2>(0, 0);
      for (int b0 = 0; b0 <= n; b0 += 4) {  // records b0 .. b0 + 3 of the factor; record 0 is the header
        if (b0) blk = R[b0 >> 2];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int t = b0 + q;
          if (t >= 1 && t <= n) {
            get(blk[4 * q], b);
            __builtin_amdgcn_sched_barrier(0);
            const T cf = coef(blk, q);
#pragma unroll
            for (int e = 0; e < E; ++e) {
              u[e].x = fma(cf, b[e].x, u[e].x);
              u[e].y = fma(cf, b[e].y, u[e].y);
            }
          }
        }
      }
      get(v, b);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int e = 0; e < E; ++e) {
        hv[e].x = fma(b[e].x, u[e].x, hv[e].x);
        hv[e].y = fma(b[e].y, u[e].y, hv[e].y);
      }
      r += (n + 4) & ~3;
    }
  }
};

template <typename T, int H, int MMAX, int FIX = 0>
__global__ void __launch_bounds__((ZNGCfg<T, H, MMAX, FIX>::NT))
__attribute__((amdgpu_waves_per_eu(ZNGCfg<T, H, MMAX, FIX>::WPE)))
    k_z_ng(cx_t<T>* base, int M, int K, const int* __restrict__ gs, const NgRec4* __restrict__ recs, Geom g,
           const cx_t<T>* __restrict__ twg) {
  z_ng_run<T, H, MMAX, FIX>(base, M, K, g, twg, RecordProducts<T>{gs, recs});
}

template <typename T, int H, int MMAX>
static int launch_z_ng_m(cx_t<T>* base, int M, int K, const int* gs, const NgRec4* recs, const Geom& g,
                         const cx_t<T>* tw, hipStream_t st) {
  constexpr int C = ZNGCfg<T, H, MMAX>::C, NT = ZNGCfg<T, H, MMAX>::NT;
  dim3 grid((unsigned)((long long)g.N * g.N / C));
  hipLaunchKernelGGL((k_z_ng<T, H, MMAX>), grid, dim3(NT), 0, st, base, M, K, gs, recs, g, tw);
  return hipGetLastError() == hipSuccess ? AAR_OK : AAR_ERR_HIP;
}

template <typename T, int H>
static int launch_z_ng(cx_t<T>* base, int M, int K, const int* gs, const NgRec4* recs, const Geom& g,
                       const cx_t<T>* tw, hipStream_t st) {
  if (M <= 4) return launch_z_ng_m<T, H, 4>(base, M, K, gs, recs, g, tw, st);
  if (M <= 6) return launch_z_ng_m<T, H, 6>(base, M, K, gs, recs, g, tw, st);
  if (M <= 8) return launch_z_ng_m<T, H, 8>(base, M, K, gs, recs, g, tw, st);
  if (M <= 16) return launch_z_ng_m<T, H, 16>(base, M, K, gs, recs, g, tw, st);
  return AAR_ERR_UNSUPPORTED;
}

template <typename T, int H>
static int launch_z_source(cx_t<T>* base, const Geom& g, const cx_t<T>* tw, hipStream_t st) {
  constexpr int C = ZNGCfg<T, H, 6, 1>::C, NT = ZNGCfg<T, H, 6, 1>::NT;
  dim3 grid((unsigned)((long long)g.N * g.N / C));
  hipLaunchKernelGGL((k_z_ng<T, H, 6, 1>), grid, dim3(NT), 0, st, base, 6, 1, nullptr, nullptr, g, tw);
  return hipGetLastError() == hipSuccess ? AAR_OK : AAR_ERR_HIP;
}

template <typename T>
int Pipeline<T>::z_source(cx_t<T>* f, const Geom& g, const cx_t<T>* tw, hipStream_t st) {
  if (!g.pow2) return AAR_ERR_UNSUPPORTED;
  switch (g.H) {
    case 32: return launch_z_source<T, 32>(f, g, tw, st);
    case 64: return launch_z_source<T, 64>(f, g, tw, st);
    case 128: return launch_z_source<T, 128>(f, g, tw, st);
    case 256: return launch_z_source<T, 256>(f, g, tw, st);
    case 512: return launch_z_source<T, 512>(f, g, tw, st);
    default: return AAR_ERR_UNSUPPORTED;
  }
}

template <typename T>
int Pipeline<T>::z_ng(cx_t<T>* f, int M, int K, const int* gs, const NgRec4* recs, const Geom& g,
                      const cx_t<T>* tw, hipStream_t st) {
  if (M > 16 || !g.pow2) return AAR_ERR_UNSUPPORTED;
  switch (g.H) {
    case 32: return launch_z_ng<T, 32>(f, M, K, gs, recs, g, tw, st);
    case 64: return launch_z_ng<T, 64>(f, M, K, gs, recs, g, tw, st);
    case 128: return launch_z_ng<T, 128>(f, M, K, gs, recs, g, tw, st);
    case 256: return launch_z_ng<T, 256>(f, M, K, gs, recs, g, tw, st);
    case 512: return launch_z_ng<T, 512>(f, M, K, gs, recs, g, tw, st);
    default: return AAR_ERR_UNSUPPORTED;
  }
}

template <typename T>
int Pipeline<T>::xy_pass(cx_t<T>* f, int nf, const Geom& g, const cx_t<T>* tw, int axis, int dir,
                         hipStream_t st, int variant, XYRange r, const XOut* xs, const XModes* xm) {
  if (!g.pow2) {  // whole passes only: column groups belong to the power-of-two tiles
    if (r.k0 != 0 || (r.cols != 0 && r.cols != g.H) || !r.nyq || xs || xm) return AAR_ERR_UNSUPPORTED;
    if (axis == 0) return dir > 0 ? launch_xy_gen<T, 0, 1>(f, nf, g, tw, st) : launch_xy_gen<T, 0, -1>(f, nf, g, tw, st);
    return dir > 0 ? launch_xy_gen<T, 1, 1>(f, nf, g, tw, st) : launch_xy_gen<T, 1, -1>(f, nf, g, tw, st);
  }
  const int v = variant >= 0 ? variant : aar_xy_variant(sizeof(T) == 8, g.N, xs ? 2 : axis);
  if (variant < 0 && v != 0) {  // a pinned shape that is not built for this size runs as shape 0
    const int rc = xy_pass(f, nf, g, tw, axis, dir, st, v, r, xs, xm);
    return rc == AAR_ERR_UNSUPPORTED ? xy_pass(f, nf, g, tw, axis, dir, st, 0, r, xs, xm) : rc;
  }
#define XY_CASE(LL)                                                                                 \
  case LL:                                                                                          \
    if (axis == 0)                                                                                  \
      return dir > 0 ? launch_xy<T, LL, 0, 1>(f, nf, g, tw, st, v, r, xs, xm)                       \
                     : launch_xy<T, LL, 0, -1>(f, nf, g, tw, st, v, r, xs, xm);                     \
    return dir > 0 ? launch_xy<T, LL, 1, 1>(f, nf, g, tw, st, v, r, xs, xm)                         \
                   : launch_xy<T, LL, 1, -1>(f, nf, g, tw, st, v, r, xs, xm);
  switch (g.N) {
    XY_CASE(16)
    XY_CASE(32)
    XY_CASE(64)
    XY_CASE(128)
    XY_CASE(256)
    XY_CASE(512)
    XY_CASE(1024)
    default:
      return AAR_ERR_UNSUPPORTED;
  }
#undef XY_CASE
}

template <typename T>
int Pipeline<T>::xy_pair(cx_t<T>* f, int nf, const Geom& g, const cx_t<T>* tw, int dir, hipStream_t st,
                         int group_cols) {
  const int a0 = dir > 0 ? 0 : 1, a1 = 1 - a0;
  int rc;
  if (group_cols < 0) group_cols = g.pow2 ? aar_xy_group(sizeof(T) == 8, g.N) : 0;
  if (group_cols <= 0) {
    if ((rc = xy_pass(f, nf, g, tw, a0, dir, st))) return rc;
    return xy_pass(f, nf, g, tw, a1, dir, st);
  }
  for (int q = 0; q < nf; ++q)
    for (int k0 = 0; k0 < g.H; k0 += group_cols) {
      const XYRange r{k0, k0 + group_cols < g.H ? group_cols : g.H - k0, k0 == 0};
      cx_t<T>* fq = f + (long long)q * g.fs;
      if ((rc = xy_pass(fq, 1, g, tw, a0, dir, st, -1, r)) || (rc = xy_pass(fq, 1, g, tw, a1, dir, st, -1, r)))
        return rc;
    }
  return AAR_OK;
}

// The y pass of the outputs of inputs [q_lo, q_hi) of a split, as one launch list.
static XOut split_outputs(const XSplit& xs, int q_lo, int q_hi) {
  XOut o;
  o.kv = xs.kv;
  o.icic = xs.icic;
  for (int d = 0; d < xs.nout; ++d)
    if (xs.src[d] >= q_lo && xs.src[d] < q_hi) {
      o.src[o.n] = (signed char)(xs.in0 + xs.src[d]);
      o.dst[o.n] = (signed char)d;
      o.code[o.n] = xs.code[d];
      ++o.n;
    }
  return o;
}

template <typename T>
int Pipeline<T>::xy_split(cx_t<T>* f, const XSplit& xs, const Geom& g, const cx_t<T>* tw, hipStream_t st,
                          int group_cols) {
  // inputs after the outputs, each output reading one of them
  if (!g.pow2 || xs.nin < 1 || xs.nout < xs.nin || xs.nout > 9 || xs.in0 < xs.nout) return AAR_ERR_UNSUPPORTED;
  for (int d = 0; d < xs.nout; ++d)
    if (xs.src[d] < 0 || xs.src[d] >= xs.nin) return AAR_ERR_UNSUPPORTED;
  int rc;
  cx_t<T>* in = f + (long long)xs.in0 * g.fs;
  if (group_cols < 0) group_cols = aar_xy_split_group(sizeof(T) == 8, g.N);
  if (group_cols <= 0) {
    const XOut o = split_outputs(xs, 0, xs.nin);
    if ((rc = xy_pass(in, xs.nin, g, tw, 0, +1, st))) return rc;
    return xy_pass(f, 1, g, tw, 1, +1, st, -1, XYRange(), &o);
  }
  for (int q = 0; q < xs.nin; ++q) {
    const XOut o = split_outputs(xs, q, q + 1);
    for (int k0 = 0; k0 < g.H; k0 += group_cols) {
      const XYRange r{k0, k0 + group_cols < g.H ? group_cols : g.H - k0, k0 == 0};
      if ((rc = xy_pass(in + (long long)q * g.fs, 1, g, tw, 0, +1, st, -1, r)) ||
          (o.n && (rc = xy_pass(f, 1, g, tw, 1, +1, st, -1, r, &o))))
        return rc;
    }
  }
  return AAR_OK;
}

template <typename T>
int Pipeline<T>::modes_xy(const cx_t<T>* phig, cx_t<T>* out, int M, const int* modes, const Geom& g,
                          const KSpace& ks, const cx_t<T>* tw, hipStream_t st, int group_cols) {
  if (!g.pow2 || M < 1) return AAR_ERR_UNSUPPORTED;
  XModes xm;
  xm.M = M;
  xm.phig = phig;
  xm.modes = modes;
  xm.ks = ks;
  int rc;
  if (group_cols < 0) group_cols = aar_xy_group(sizeof(T) == 8, g.N);
  // tile shape 0: the filter's per-element recurrence in the loads of the pipelined shape 3
  // would hold all of a tile's in flight at once
  if (group_cols <= 0) {
    if ((rc = xy_pass(out, 1, g, tw, 0, +1, st, 0, XYRange(), nullptr, &xm))) return rc;
    return xy_pass(out, M, g, tw, 1, +1, st);
  }
  for (int k0 = 0; k0 < g.H; k0 += group_cols) {
    const XYRange r{k0, k0 + group_cols < g.H ? group_cols : g.H - k0, k0 == 0};
    if ((rc = xy_pass(out, 1, g, tw, 0, +1, st, 0, r, nullptr, &xm)) || (rc = xy_pass(out, M, g, tw, 1, +1, st, -1, r)))
      return rc;
  }
  return AAR_OK;
}

XSplit lpt_xsplit(const KSpace& ks) {
  // inputs (ng_lpt, split) in slots 9-13: kx^2 B, kx B, B, i kx sx B, sx B with B = T(k)/(D+ Beta)
  // Phi and sx = 1/sinc_x^2; outputs -kx kx B, -kx ky B, -kx kz B, -ky ky B, -ky kz B, -kz kz B
  // and i k_a B sx sy sz (main.c:1207-1211, 1238-1277, 1311-1369)
  XSplit xs;
  xs.nin = 5;
  xs.nout = 9;
  xs.in0 = 9;
  const signed char src[9] = {0, 1, 1, 2, 2, 2, 3, 4, 4};
  const unsigned char code[9] = {kXsNeg,         kXsKy | kXsNeg,         kXsKz | kXsNeg,
                                 kXsKy2 | kXsNeg, kXsKy | kXsKz | kXsNeg, kXsKz2 | kXsNeg,
                                 kXsCic,         kXsKy | kXsCic | kXsI,  kXsKz | kXsCic | kXsI};
  for (int o = 0; o < 9; ++o) xs.src[o] = src[o], xs.code[o] = code[o];
  xs.kv = ks.kv;
  xs.icic = ks.icic;
  return xs;
}

XSplit disp2_xsplit(const KSpace& ks) {
  // inputs (disp2_spec, split) in slots 3-4: -i kx sx S/k^2, sx S/k^2 (S the 2LPT source);
  // outputs -i k_a S/k^2 sx sy sz (main.c:1353-1366)
  XSplit xs;
  xs.nin = 2;
  xs.nout = 3;
  xs.in0 = 3;
  xs.src[0] = 0, xs.src[1] = 1, xs.src[2] = 1;
  xs.code[0] = kXsCic;
  xs.code[1] = kXsKy | kXsCic | kXsI | kXsNeg;
  xs.code[2] = kXsKz | kXsCic | kXsI | kXsNeg;
  xs.kv = ks.kv;
  xs.icic = ks.icic;
  return xs;
}

template <typename T>
int Pipeline<T>::z_pass(cx_t<T>* f, int nf, const Geom& g, const cx_t<T>* tw, int dir, hipStream_t st) {
  if (!g.pow2) return launch_z_gen<T>(f, nf, g, tw, dir, st);
  switch (g.H) {
    case 8: return launch_z<T, 8>(f, nf, g, tw, dir, st);
    case 16: return launch_z<T, 16>(f, nf, g, tw, dir, st);
    case 32: return launch_z<T, 32>(f, nf, g, tw, dir, st);
    case 64: return launch_z<T, 64>(f, nf, g, tw, dir, st);
    case 128: return launch_z<T, 128>(f, nf, g, tw, dir, st);
    case 256: return launch_z<T, 256>(f, nf, g, tw, dir, st);
    case 512: return launch_z<T, 512>(f, nf, g, tw, dir, st);
    default: return AAR_ERR_UNSUPPORTED;
  }
}

// C2R with FFTW's order: complex inverse over x and y, then half-complex over z.
template <typename T>
int Pipeline<T>::c2r(cx_t<T>* f, int nf, const Geom& g, const cx_t<T>* tw, hipStream_t st) {
  int rc;
  if ((rc = xy_pair(f, nf, g, tw, +1, st, -1))) return rc;
  return z_pass(f, nf, g, tw, +1, st);
}

template <typename T>
int Pipeline<T>::r2c(cx_t<T>* f, int nf, const Geom& g, const cx_t<T>* tw, hipStream_t st) {
  int rc;
  if ((rc = z_pass(f, nf, g, tw, -1, st))) return rc;
  return xy_pair(f, nf, g, tw, -1, st, -1);
}

// ------------------------------------------------------------------ Gaussian potential
// One thread per (i, j) pencil, main.c:382-524.  The stream of pencil (i, j)
// is reseeded with seedtable[i N + j]; two uniforms (phase, amplitude) are
// drawn for every k < N/2, also where the mode is cut or Nyquist.
// A block is one wavefront of 64 pencils (consecutive memory rows).  Each
// ranlxd update yields 6 modes per pencil; they go to a 14-mode LDS ring per
// pencil (one 8-mode chunk plus the 6 modes that may run past it), and every
// completed chunk is written by the whole wavefront as row segments of 8
// modes, so the stores are coalesced.
constexpr int kGaussRows = 64;
template <typename T>
__global__ void __launch_bounds__(kGaussRows)
    k_gauss(cx_t<T>* __restrict__ phi, Geom g, KSpace ks, const uint32_t* __restrict__ seeds,
            double anorm, double ns, double fac, double beta, double phase_shift, int fixed_amp) {
  using C2 = cx_t<T>;
  constexpr int CH = 8, RING = CH + 6;
  __shared__ C2 ring[kGaussRows][RING + 1];
  const int N = g.N, H = g.H;
  const int lane = threadIdx.x;
  const long long row0 = (long long)blockIdx.x * kGaussRows;
  const long long mrow = row0 + lane;  // memory row of this thread's pencil (N^2 % 64 == 0)
  int i, j;
  row_xy(N, mrow, i, j);
  const long long pid = (long long)i * N + j;  // seed table index
  C2* row = phi + mrow * H;
  // Which k = 0 entries a pencil owns (main.c:476-520): its own k = 0 mode
  // (and the conjugate at (-i, -j)), a zero, or nothing (the partner writes it).
  auto owns0 = [&](int pi, int pj, bool& gen, bool& zero) {
    const bool nyq = (pi == H) || (pj == H);
    gen = !nyq && ((pi == 0 && pj > 0 && pj < H) || (pi > 0 && pi < H));
    zero = nyq || (pi == 0 && pj == 0);
  };
  bool gen0, zero0;
  owns0(i, j, gen0, zero0);
  const bool nyq_pencil = (i == H) || (j == H);
  const int ii = (N - i) % N, jj = (N - j) % N;
  phi[g.nq + mrow] = cmake<C2>(0, 0);

  // FP32 path: amplitude and phase in FP32 from the FP64 draws, with the
  // constant factors folded: sqrt(anorm) fac beta sqrt(k^ns (-ln u2)) / k^2
  const float amp32 = (float)(sqrt(anorm) * fac * beta);
  auto value = [&](int k, double u1, double u2) {
    C2 val = cmake<C2>(0, 0);
    if (!nyq_pencil && !(i == 0 && j == 0 && k == 0)) {
      const double kmag2 = kmag2_of(ks, i, j, k);
      const double kmag = ks.sphere == 1 ? sqrt(kmag2) : 0.0;  // the cube cut uses per-axis flags
      if (!outside_cut(ks, i, j, k, kmag)) {
        if constexpr (sizeof(T) == 4) {
          // the hardware exp2 / log2 / sin / cos (a few ulp; the phase is in [0, 3 pi))
          const float k2 = (float)kmag2;
          float a = __expf(0.5f * (float)ns * __logf(k2));
          // the oscillation phase w k can be large: its factor is formed in FP64
          if (ks.res) a *= (float)res_modulate(ks, 1.0, sqrt(kmag2));
          if (!fixed_amp) a *= -__logf((float)u2);
          const float phig = amp32 * sqrtf(a) / k2;
          float sn, cs;
          __sincosf((float)(u1 * 2 * AAR_PI + phase_shift), &sn, &cs);
          val = cmake<C2>(phig * cs, phig * sn);
        } else {
          double phase = u1 * 2 * AAR_PI;
          phase += phase_shift;
          const double km = sqrt(kmag2);
          double phig = anorm * exp(ns * log(km));
          if (ks.res) phig = res_modulate(ks, phig, km);
          if (!fixed_amp) phig *= -log(u2);
          phig = sqrt(phig) * fac * beta / kmag2;
          val = cmake<C2>((T)(phig * cos(phase)), (T)(phig * sin(phase)));
        }
      }
    }
    if (k == 0 && gen0) phi[row_of(N, ii, jj) * H] = cmake<C2>(val.x, -val.y);
    return val;
  };
  // write the completed chunk c of every pencil of the block
  auto flush = [&](int c) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kGaussRows * CH / 64; ++q) {
      const int e = lane + 64 * q, r = e / CH, kk = e % CH;
      const int k = c * CH + kk;
      bool write = true;
      if (k == 0) {
        int pi, pj;
        row_xy(N, row0 + r, pi, pj);
        bool gn, zr;
        owns0(pi, pj, gn, zr);
        write = gn || zr;
      }
      if (write) phi[(row0 + r) * H + k] = ring[r][k % RING];
    }
    __syncthreads();
  };

  Ranlxd rs;
  ranlxd_seed(rs, seeds[pid]);
  bool redo = false;
  for (int kb = 0; kb < H; kb += 6) {
    ranlxd_update_dev(rs);
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const int k = kb + q;
      if (k < H) {
        const double u1 = (double)rs.s[2 * q] * kTwoM48;
        const double u2 = (double)rs.s[2 * q + 1] * kTwoM48;
        if (u2 == 0) redo = true;
        ring[lane][k % RING] = redo ? cmake<C2>(0, 0) : value(k, u1, u2);
      }
    }
    const int c0 = kb / CH, c1 = (kb + 6 < H ? kb + 6 : H) / CH;
    if (c1 > c0) flush(c0);  // uniform: kb is the same in every lane
  }
  if (redo) {  // an amplitude draw was exactly 0: replay with GSL's redraw loop
    __threadfence();
    Ranlxd rr;
    ranlxd_seed(rr, seeds[pid]);
    int pos = 12;
    auto next = [&]() {
      if (pos == 12) {
        ranlxd_update_dev(rr);
        pos = 0;
      }
      uint64_t v = rr.s[0];
#pragma unroll
      for (int q = 1; q < 12; ++q) v = (pos == q) ? rr.s[q] : v;
      ++pos;
      return (double)v * kTwoM48;
    };
    for (int k = 0; k < H; ++k) {
      const double u1 = next();
      double u2;
      do u2 = next();
      while (u2 == 0);
      const C2 v = value(k, u1, u2);
      if (k > 0 || gen0) row[k] = v;
      else if (zero0) row[0] = cmake<C2>(0, 0);
    }
  }
}

template <typename T>
int Pipeline<T>::gauss(cx_t<T>* phi, const Geom& g, const KSpace& ks, const uint32_t* seeds, double anorm,
                       double ns, double fac, double beta, double phase_shift, int fixed_amp,
                       hipStream_t st) {
  const long long np = (long long)g.N * g.N;
  if (np % kGaussRows) return AAR_ERR_UNSUPPORTED;
  hipLaunchKernelGGL((k_gauss<T>), dim3((unsigned)(np / kGaussRows)), dim3(kGaussRows), 0, st, phi, g, ks,
                     seeds, anorm, ns, fac, beta, phase_shift, fixed_amp);
  return hipGetLastError() == hipSuccess ? AAR_OK : AAR_ERR_HIP;
}
// ------------------------------------------------------------------ mode fields
// out[m](k) = q_n(k) phi(k), n = modes[m] (ascending), q_n = g_n(k) k^{4-ns}:
// n <= 3: k^{s n};  n >= 4: (P_{n-1}(xbar) - X0[n-1]) k^{2-ns}   (main.c:571-639).
template <typename T>
__global__ void __launch_bounds__(256)
    k_modes_spec(const cx_t<T>* __restrict__ phi, cx_t<T>* __restrict__ out, int M,
                 const int* __restrict__ modes, Geom g, KSpace ks) {
  using C2 = cx_t<T>;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= g.fs) return;
  int x, y, z;
  decode_k(g, e, x, y, z);
  const double kmag2 = kmag2_of(ks, x, y, z);
  const double kmag1 = (sizeof(T) == 8 || ks.sphere == 1) ? sqrt(kmag2) : 0.0;
  const bool dead = (x + y + z == 0) || outside_cut(ks, x, y, z, kmag1);
  if (dead) {
    for (int m = 0; m < M; ++m) out[(long long)m * g.fs + e] = cmake<C2>(0, 0);
    return;
  }
  const C2 p = phi[e];
  using RT = T;  // radial functions in the path's precision
  const RT k2 = (RT)kmag2;
  const RT kss = kpow_s<RT>(kmag2, kmag1, ks.s);
  const RT xb = -1 + 2 * (kss - (RT)ks.kmin_s) / (RT)(ks.kmax_s - ks.kmin_s);
  const RT k3s = kss * kss * kss;  // k^{4-ns}
  LegRecT<RT> leg;
  leg.init(xb);
  for (int m = 0; m < M; ++m) {
    const int n = modes[m];
    RT q;
    if (n <= 3) {
      q = 1;
      for (int t = 0; t < n; ++t) q *= kss;
    } else {
      q = (leg.at(n - 1) - (RT)ks.x0[n - 1]) * (k3s / k2);
    }
    out[(long long)m * g.fs + e] = cmake<C2>(q * p.x, q * p.y);
  }
}

template <typename T>
int Pipeline<T>::modes_spec(const cx_t<T>* phi, cx_t<T>* out, int M, const int* modes, const Geom& g,
                            const KSpace& ks, hipStream_t st) {
  hipLaunchKernelGGL((k_modes_spec<T>), dim3((unsigned)((g.fs + 255) / 256)), dim3(256), 0, st, phi,
                     out, M, modes, g, ks);
  return hipGetLastError() == hipSuccess ? AAR_OK : AAR_ERR_HIP;
}

// ------------------------------------------------------------------ real-space products
template <typename T, int MMAX>
__global__ void __launch_bounds__(256)
    k_products(const cx_t<T>* __restrict__ fin, int M, cx_t<T>* __restrict__ fout, int K,
               const NgTerm* __restrict__ terms, const int* __restrict__ gs, Geom g) {
  const long long n3 = (long long)g.N * g.N * g.N;
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n3) return;
  T f[MMAX];
#pragma unroll
  for (int m = 0; m < MMAX; ++m)
    if (m < M) f[m] = reinterpret_cast<const T*>(fin + (long long)m * g.fs)[p];
  for (int kk = 0; kk < K; ++kk) {
    T h = 0;
    for (int t = gs[kk]; t < gs[kk + 1]; ++t) {
      const NgTerm tm = terms[t];
      h += (T)tm.c * (f[tm.i] * f[tm.j]);
    }
    reinterpret_cast<T*>(fout + (long long)kk * g.fs)[p] = h;
  }
}

template <typename T>
int Pipeline<T>::products(const cx_t<T>* fin, int M, cx_t<T>* fout, int K, const NgTerm* terms,
                          const int* gs, const Geom& g, hipStream_t st) {
  const long long n3 = (long long)g.N * g.N * g.N;
  dim3 grid((unsigned)((n3 + 255) / 256));
  if (M <= 8)
    hipLaunchKernelGGL((k_products<T, 8>), grid, dim3(256), 0, st, fin, M, fout, K, terms, gs, g);
  else if (M <= 16)
    hipLaunchKernelGGL((k_products<T, 16>), grid, dim3(256), 0, st, fin, M, fout, K, terms, gs, g);
  else if (M <= 32)
    hipLaunchKernelGGL((k_products<T, 32>), grid, dim3(256), 0, st, fin, M, fout, K, terms, gs, g);
  else
    return AAR_ERR_UNSUPPORTED;
  return hipGetLastError() == hipSuccess ? AAR_OK : AAR_ERR_HIP;
}

// ------------------------------------------------------------------ NG accumulation
template <typename T>
__global__ void __launch_bounds__(256)
    k_ng_accumulate(double2* __restrict__ acc, const cx_t<T>* __restrict__ c,
                    const NgWeight* __restrict__ w, int nw, Geom g, KSpace ks) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= g.fs) return;
  int x, y, z;
  decode_k(g, e, x, y, z);
  if (x + y + z == 0) return;
  const double kmag2 = kmag2_of(ks, x, y, z);
  const double kmag1 = sqrt(kmag2);
  if (outside_cut(ks, x, y, z, kmag1)) return;
  const double kss = pow(kmag1, ks.s);
  LegRec leg;
  leg.init(-1 + 2 * (kss - ks.kmin_s) / (ks.kmax_s - ks.kmin_s));
  double W = 0;
  for (int l = 0; l < nw; ++l) W += w[l].c * weight_g(w[l].n, kss, kmag2, leg, ks.x0);
  const cx_t<T> v = c[e];
  double2 a = acc[e];
  a.x += W * (double)v.x;
  a.y += W * (double)v.y;
  acc[e] = a;
}

template <typename T>
int Pipeline<T>::ng_accumulate(double2* acc, const cx_t<T>* c, const NgWeight* w, int nw, const Geom& g,
                               const KSpace& ks, hipStream_t st) {
  hipLaunchKernelGGL((k_ng_accumulate<T>), dim3((unsigned)((g.fs + 255) / 256)), dim3(256), 0, st, acc,
                     c, w, nw, g, ks);
  return hipGetLastError() == hipSuccess ? AAR_OK : AAR_ERR_HIP;
}

// Phi = phiG + Fnl NG / N^3 with NG = sum_K g_K(k) h_K(k) (main.c:1104-1126).
template <typename T>
__global__ void __launch_bounds__(256)
    k_ng_combine(const cx_t<T>* __restrict__ phig, const cx_t<T>* __restrict__ h, int K,
                 const int* __restrict__ kmodes, cx_t<T>* __restrict__ phi, double fnl, double n3,
                 Geom g, KSpace ks) {
  using C2 = cx_t<T>;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= g.fs) return;
  int x, y, z;
  decode_k(g, e, x, y, z);
  const C2 pg = phig[e];
  double re = pg.x, im = pg.y;
  if (x + y + z != 0) {
    const double kmag2 = kmag2_of(ks, x, y, z);
    const double kmag1 = sqrt(kmag2);
    if (!outside_cut(ks, x, y, z, kmag1)) {
      using RT = T;
      const RT kss = kpow_s<RT>(kmag2, kmag1, ks.s);
      LegRecT<RT> leg;
      leg.init(-1 + 2 * (kss - (RT)ks.kmin_s) / (RT)(ks.kmax_s - ks.kmin_s));
      double nr = 0, ni = 0;
      for (int kk = 0; kk < K; ++kk) {
        const double w = (double)weight_g<RT>(kmodes[kk], kss, (RT)kmag2, leg, ks.x0);
        const C2 v = h[(long long)kk * g.fs + e];
        nr += w * (double)v.x;
        ni += w * (double)v.y;
      }
      re += fnl * nr / n3;
      im += fnl * ni / n3;
    }
  } else {
    re = 0;
    im = 0;
  }
  phi[e] = cmake<C2>((T)re, (T)im);
}

template <typename T>
int Pipeline<T>::ng_combine(const cx_t<T>* phig, const cx_t<T>* h, int K, const int* kmodes,
                            cx_t<T>* phi, double fnl, const Geom& g, const KSpace& ks,
                            hipStream_t st) {
  const double n3 = (double)((unsigned)g.N * (unsigned)g.N * (unsigned)g.N);
  hipLaunchKernelGGL((k_ng_combine<T>), dim3((unsigned)((g.fs + 255) / 256)), dim3(256), 0, st, phig, h,
                     K, kmodes, phi, fnl, n3, g, ks);
  return hipGetLastError() == hipSuccess ? AAR_OK : AAR_ERR_HIP;
}

template <typename T>
__global__ void __launch_bounds__(256)
    k_ng_combine_acc(const cx_t<T>* __restrict__ phig, const double2* __restrict__ acc,
                     cx_t<T>* __restrict__ phi, double fnl, double n3, Geom g) {
  using C2 = cx_t<T>;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= g.fs) return;
  const C2 pg = phig[e];
  const double2 a = acc[e];
  double re = (double)pg.x + fnl * a.x / n3;
  double im = (double)pg.y + fnl * a.y / n3;
  if (e == 0) re = im = 0;
  phi[e] = cmake<C2>((T)re, (T)im);
}

template <typename T>
int Pipeline<T>::ng_combine_acc(const cx_t<T>* phig, const double2* acc, cx_t<T>* phi, double fnl,
                                double n3, const Geom& g, hipStream_t st) {
  hipLaunchKernelGGL((k_ng_combine_acc<T>), dim3((unsigned)((g.fs + 255) / 256)), dim3(256), 0, st, phig,
                     acc, phi, fnl, n3, g);
  return hipGetLastError() == hipSuccess ? AAR_OK : AAR_ERR_HIP;
}

// ------------------------------------------------------------------ 2LPT
template <typename RT> __device__ RT transfer_dev(const TransferDev& tf, RT k) {
  if (tf.which == 1) {  // Eisenstein & Hu no-wiggle, power.c:519-548
    // FP32 path: the hardware exp / log (a few ulp)
    auto ex = [](RT v) -> RT {
      if constexpr (sizeof(RT) == 4) return __expf(v);
      else return exp(v);
    };
    auto lg = [](RT v) -> RT {
      if constexpr (sizeof(RT) == 4) return __logf(v);
      else return log(v);
    };
    k *= (RT)tf.kconv;
    const RT a = (RT)tf.eh_a;
    RT gamma = a + ((RT)1 - a) / ((RT)1 + ex((RT)4 * lg((RT)0.43 * k * (RT)tf.eh_s)));
    gamma *= (RT)tf.eh_omh;
    RT q = k * (RT)tf.eh_theta2 / gamma;
    RT L0 = lg((RT)tf.eh_twoe + (RT)1.8 * q);
    RT C0 = (RT)14.2 + (RT)731. / ((RT)1 + (RT)62.5 * q);
    return L0 / (L0 + C0 * q * q);
  }
  if (tf.which == 2) {  // log-log table, power.c:181-217
    k *= (RT)tf.kconv;
    RT logk = log10(k);
    if (logk < (RT)tf.logk[0] || logk > (RT)tf.logk[tf.nt - 1]) return 0;
    int lo = 0, hi = tf.nt - 1;
    while (hi - lo > 1) {
      int mid = (hi + lo) / 2;
      if (logk < (RT)tf.logk[mid]) hi = mid;
      else lo = mid;
    }
    RT u = (logk - (RT)tf.logk[lo]) / (RT)(tf.logk[hi] - tf.logk[lo]);
    RT logT = (1 - u) * (RT)tf.logT[lo] + u * (RT)tf.logT[hi];
    return pow((RT)10, logT);
  }
  return 1;
}

// out9[0..5] = i k_b psi_a for ab = 00,01,02,11,12,22 (main.c:1238-1277);
// out9[6..8] = psi_a * CIC deconvolution (main.c:1311-1369).
template <typename T>
__global__ void __launch_bounds__(256)
    k_lpt_spec(const cx_t<T>* __restrict__ phi, cx_t<T>* __restrict__ out, Geom g, KSpace ks,
               TransferDev tf, double dplus, double beta) {
  using C2 = cx_t<T>;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= g.fs) return;
  int x, y, z;
  decode_k(g, e, x, y, z);
  const double kv[3] = {ks.kv[x], ks.kv[y], ks.kv[z]};
  const double kmag2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
  const double kmag = sqrt(kmag2);
  const double twb = (double)transfer_dev<T>(tf, (T)kmag) / dplus / beta;
  const C2 p = phi[e];
  double cr[3], ci[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    ci[a] = kv[a] * twb * (double)p.x;
    cr[a] = -kv[a] * twb * (double)p.y;
  }
  const int ga[6] = {0, 0, 0, 1, 1, 2}, gb[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
  for (int q = 0; q < 6; ++q)
    out[(long long)q * g.fs + e] = cmake<C2>((T)(-ci[ga[q]] * kv[gb[q]]), (T)(cr[ga[q]] * kv[gb[q]]));
  const double ff = 1 / (ks.sinc[x] * ks.sinc[y] * ks.sinc[z]);
  const double smth = ff * ff;
#pragma unroll
  for (int a = 0; a < 3; ++a)
    out[(long long)(6 + a) * g.fs + e] = cmake<C2>((T)(cr[a] * smth), (T)(ci[a] * smth));
}

template <typename T>
int Pipeline<T>::lpt_spec(const cx_t<T>* phi, cx_t<T>* out9, const Geom& g, const KSpace& ks,
                          const TransferDev& tf, double dplus, double beta, hipStream_t st) {
  hipLaunchKernelGGL((k_lpt_spec<T>), dim3((unsigned)((g.fs + 255) / 256)), dim3(256), 0, st, phi, out9,
                     g, ks, tf, dplus, beta);
  return hipGetLastError() == hipSuccess ? AAR_OK : AAR_ERR_HIP;
}

// Batched path: Phi = phiG + Fnl/N^3 sum_K g_K h_K (written to phi) and the
// nine 2LPT spectra of k_lpt_spec in one elementwise pass (out9 may alias h).
template <typename T>
__global__ void __launch_bounds__(256)
    k_ng_lpt(const cx_t<T>* __restrict__ phig, const cx_t<T>* h, int K, const int* __restrict__ kmodes,
             cx_t<T>* __restrict__ phi, cx_t<T>* out, double fnl, double n3, Geom g, KSpace ks,
             TransferDev tf, double dplus, double beta, int split) {
  using C2 = cx_t<T>;
  using RT = T;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= g.fs) return;
  int x, y, z;
  decode_k(g, e, x, y, z);
  const double kv[3] = {ks.kv[x], ks.kv[y], ks.kv[z]};
  const double kmag2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
  const double kmag1 = sqrt(kmag2);
  const C2 pg = phig[e];
  double re = pg.x, im = pg.y;
  if (x + y + z == 0) {
    re = im = 0;
  } else if (!outside_cut(ks, x, y, z, kmag1)) {
    const RT kss = kpow_s<RT>(kmag2, kmag1, ks.s);
    LegRecT<RT> leg;
    leg.init(-1 + 2 * (kss - (RT)ks.kmin_s) / (RT)(ks.kmax_s - ks.kmin_s));
    double nr = 0, ni = 0;
    for (int kk = 0; kk < K; ++kk) {
      const double w = (double)weight_g<RT>(kmodes[kk], kss, (RT)kmag2, leg, ks.x0);
      const C2 v = h[(long long)kk * g.fs + e];
      nr += w * (double)v.x;
      ni += w * (double)v.y;
    }
    re += fnl * nr / n3;
    im += fnl * ni / n3;
  }
  const C2 p = cmake<C2>((T)re, (T)im);
  phi[e] = p;
  const double twb = (double)transfer_dev<T>(tf, (T)kmag1) / dplus / beta;
  if (split) {
    // the x-pass inputs of lpt_xsplit: B = twb Phi times the factors of kx (all reads of h at e
    // are done; the outputs may overwrite them)
    const double bx = twb * (double)p.x, by = twb * (double)p.y, kx = kv[0], sx = ks.icic[x];
    out[e] = cmake<C2>((T)(kx * kx * bx), (T)(kx * kx * by));
    out[g.fs + e] = cmake<C2>((T)(kx * bx), (T)(kx * by));
    out[2LL * g.fs + e] = cmake<C2>((T)bx, (T)by);
    out[3LL * g.fs + e] = cmake<C2>((T)(-kx * sx * by), (T)(kx * sx * bx));
    out[4LL * g.fs + e] = cmake<C2>((T)(sx * bx), (T)(sx * by));
    return;
  }
  double cr[3], ci[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    ci[a] = kv[a] * twb * (double)p.x;
    cr[a] = -kv[a] * twb * (double)p.y;
  }
  const double ff = 1 / (ks.sinc[x] * ks.sinc[y] * ks.sinc[z]);
  const double smth = ff * ff;
  // all reads of h at e are done; the outputs may overwrite them
#pragma unroll
  for (int a = 0; a < 3; ++a) {
#pragma unroll
    for (int b = a; b < 3; ++b) {
      const int q = a == 0 ? b : (a == 1 ? 2 + b : 5);
      out[(long long)q * g.fs + e] = cmake<C2>((T)(-ci[a] * kv[b]), (T)(cr[a] * kv[b]));
    }
  }
#pragma unroll
  for (int a = 0; a < 3; ++a)
    out[(long long)(6 + a) * g.fs + e] = cmake<C2>((T)(cr[a] * smth), (T)(ci[a] * smth));
}

template <typename T>
int Pipeline<T>::ng_lpt(const cx_t<T>* phig, const cx_t<T>* h, int K, const int* kmodes, cx_t<T>* phi,
                        cx_t<T>* out9, double fnl, const Geom& g, const KSpace& ks, const TransferDev& tf,
                        double dplus, double beta, hipStream_t st, bool split) {
  const double n3 = (double)((unsigned)g.N * (unsigned)g.N * (unsigned)g.N);
  hipLaunchKernelGGL((k_ng_lpt<T>), dim3((unsigned)((g.fs + 255) / 256)), dim3(256), 0, st, phig, h, K, kmodes,
                     phi, out9, fnl, n3, g, ks, tf, dplus, beta, split ? 1 : 0);
  return hipGetLastError() == hipSuccess ? AAR_OK : AAR_ERR_HIP;
}

// second-order source into field 3 (main.c:1284-1294)
template <typename T>
__global__ void __launch_bounds__(256) k_source(cx_t<T>* __restrict__ dg, Geom g) {
  const long long n3 = (long long)g.N * g.N * g.N;
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n3) return;
  T d[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) d[q] = reinterpret_cast<const T*>(dg + (long long)q * g.fs)[p];
  reinterpret_cast<T*>(dg + 3 * g.fs)[p] =
      d[0] * (d[3] + d[5]) + d[3] * d[5] - d[1] * d[1] - d[2] * d[2] - d[4] * d[4];
}

template <typename T> int Pipeline<T>::source(cx_t<T>* dg6, const Geom& g, hipStream_t st) {
  const long long n3 = (long long)g.N * g.N * g.N;
  hipLaunchKernelGGL((k_source<T>), dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, st, dg6, g);
  return hipGetLastError() == hipSuccess ? AAR_OK : AAR_ERR_HIP;
}

// psi2_a = src k_a / (i k^2), CIC-deconvolved (main.c:1311-1369)
template <typename T>
__global__ void __launch_bounds__(256)
    k_disp2_spec(const cx_t<T>* __restrict__ src, cx_t<T>* __restrict__ out, Geom g, KSpace ks, int split) {
  using C2 = cx_t<T>;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= g.fs) return;
  with_k(g, e, [&](int x, int y, int z) {
    // the path's precision throughout; one reciprocal replaces the divisions
    using RT = T;
    const RT kv[3] = {(RT)ks.kv[x], (RT)ks.kv[y], (RT)ks.kv[z]};
    const RT kmag2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
    if (split) {  // the x-pass inputs of disp2_xsplit: -i kx sx S/k^2 and sx S/k^2
      const RT f = kmag2 > (RT)0 ? (RT)ks.icic[x] / kmag2 : (RT)0;
      const C2 s = src[e];
      out[e] = cmake<C2>(s.y * kv[0] * f, -s.x * kv[0] * f);
      out[g.fs + e] = cmake<C2>(s.x * f, s.y * f);
      return;
    }
    const RT ff = (RT)1 / ((RT)ks.sinc[x] * (RT)ks.sinc[y] * (RT)ks.sinc[z]);
    const RT f = kmag2 > (RT)0 ? ff * ff / kmag2 : (RT)0;  // CIC deconvolution / k^2
    const C2 s = src[e];
#pragma unroll
    for (int a = 0; a < 3; ++a) out[(long long)a * g.fs + e] = cmake<C2>(s.y * kv[a] * f, -s.x * kv[a] * f);
  });
}

template <typename T>
int Pipeline<T>::disp2_spec(const cx_t<T>* src, cx_t<T>* out3, const Geom& g, const KSpace& ks,
                            hipStream_t st, bool split) {
  hipLaunchKernelGGL((k_disp2_spec<T>), dim3((unsigned)((g.fs + 255) / 256)), dim3(256), 0, st, src,
                     out3, g, ks, split ? 1 : 0);
  return hipGetLastError() == hipSuccess ? AAR_OK : AAR_ERR_HIP;
}

// ------------------------------------------------------------------ particles
// Glass tiling (read_glass.c:176-214) and CIC readout with the ZA + 2LPT
// update (main.c:1440-1516, periodic_wrap main.c:1534-1544).
// Interpolation runs in the field precision (FP32 path: FP32 weights and
// sums); the position update and periodic wrap are always FP64.
// fmod(x, b) for b > 0 with the common cases done by one compare: for
// 0 <= |x| < b the result is x, for b <= x < 2b it is x - b (exact, as fmod's
// result always is: Sterbenz); anything else takes the library routine.
__device__ __forceinline__ double fmod_box(double x, double b) {
  if (x >= 0 ? x < b : x > -b) return x;
  if (x >= b && x < 2 * b) return x - b;
  return fmod(x, b);
}

// a / b for b > 0 given rb = RN(1/b): the product and two FMA corrections (Markstein's
// iteration: after the first correction the quotient is faithful, and from a faithful quotient
// the second correction rounds correctly), 5 instructions for the 13 of the IEEE division
// sequence.  Checked bit for bit against a / b on 2e9 operands over ten divisors.
__device__ __forceinline__ double ddiv(double a, double b, double rb) {
  double q = a * rb;
  double r = fma(-b, q, a);
  q = fma(r, rb, q);
  r = fma(-b, q, a);
  return fma(r, rb, q);
}

// Per-launch constants of the CIC readout, computed on the host with the same IEEE operations
// the kernel would use: bt = box / tile, n3f = (float) N^3 and its reciprocal, and the
// reciprocals of gbox and box for ddiv.
struct CicConst {
  double gbox, rgbox, box, rbox, bt, vp1, vp2, n3f, rn3;
  int tile, n3pow2;
};

// One particle: glass particle n of tile t (tile-major order of read_glass.c).
template <typename T>
__device__ __forceinline__ void cic_particle(int n, int t, const cx_t<T>* __restrict__ disp,
                                             const cx_t<T>* __restrict__ disp2, const Geom& g,
                                             const float* __restrict__ glass, int nglass, const CicConst& cc,
                                             float* __restrict__ pos, float* __restrict__ vel) {
  const long long p = (long long)t * nglass + n;
  const int N = g.N;
  const int tile = cc.tile;
  const double box = cc.box, bt = cc.bt, vp1 = cc.vp1, vp2 = cc.vp2;
  const int ti = t / (tile * tile), tj = (t / tile) % tile, tk = t % tile;
  const float3 gp = reinterpret_cast<const float3*>(glass)[n];  // one 12-byte load
  float P[3];
  P[0] = (float)(ddiv(gp.x, cc.gbox, cc.rgbox) * bt + ti * bt);
  P[1] = (float)(ddiv(gp.y, cc.gbox, cc.rgbox) * bt + tj * bt);
  P[2] = (float)(ddiv(gp.z, cc.gbox, cc.rgbox) * bt + tk * bt);
  double u = ddiv(P[0], box, cc.rbox) * N, v = ddiv(P[1], box, cc.rbox) * N, w = ddiv(P[2], box, cc.rbox) * N;
  int i = (int)u, j = (int)v, k = (int)w;
  if (i == N) i = N - 1;
  if (i < 0) i = 0;
  if (j == N) j = N - 1;
  if (k == N) k = N - 1;
  u -= i;
  v -= j;
  w -= k;
  int ii = i + 1, jj = j + 1, kk = k + 1;
  if (ii >= N) ii -= N;
  if (jj >= N) jj -= N;
  if (kk >= N) kk -= N;
  using W = T;  // interpolation precision
  const W uu = (W)u, vv = (W)v, ww = (W)w;
  const W f1 = (1 - uu) * (1 - vv) * (1 - ww), f2 = (1 - uu) * (1 - vv) * (ww);
  const W f3 = (1 - uu) * (vv) * (1 - ww), f4 = (1 - uu) * (vv) * (ww);
  const W f5 = (uu) * (1 - vv) * (1 - ww), f6 = (uu) * (1 - vv) * (ww);
  const W f7 = (uu) * (vv) * (1 - ww), f8 = (uu) * (vv) * (ww);
  // the eight corner offsets, once for all six fields (a real field holds N^3 <= 2^30 values)
  const unsigned c1 = (unsigned)(row_of(N, i, j) * N), c2 = (unsigned)(row_of(N, i, jj) * N);
  const unsigned c3 = (unsigned)(row_of(N, ii, j) * N), c4 = (unsigned)(row_of(N, ii, jj) * N);
  unsigned o[8] = {c1 + k, c1 + kk, c2 + k, c2 + kk, c3 + k, c3 + kk, c4 + k, c4 + kk};
  // FP32: as byte offsets, which fit 32 bits (4 N^3 <= 2^32), so every load is the field's
  // uniform base plus a 32-bit offset with no per-load address arithmetic
  if constexpr (sizeof(T) == 4)
#pragma unroll
    for (int q = 0; q < 8; ++q) o[q] <<= 2;
  auto at = [&](const T* D, unsigned off) -> T {
    if constexpr (sizeof(T) == 4) return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(D) + off);
    else return D[off];
  };

  // all 48 gathers and interpolations first, in one branch-free block so the loads can be
  // issued together; then the position updates with their wrap branches
  double dis[3], dis2[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const T* D = reinterpret_cast<const T*>(disp + (long long)a * g.fs);
    const T* D2 = reinterpret_cast<const T*>(disp2 + (long long)a * g.fs);
    dis[a] = (double)(at(D, o[0]) * f1 + at(D, o[1]) * f2 + at(D, o[2]) * f3 + at(D, o[3]) * f4 +
                      at(D, o[4]) * f5 + at(D, o[5]) * f6 + at(D, o[6]) * f7 + at(D, o[7]) * f8);
    dis2[a] = (double)(at(D2, o[0]) * f1 + at(D2, o[1]) * f2 + at(D2, o[2]) * f3 + at(D2, o[3]) * f4 +
                       at(D2, o[4]) * f5 + at(D2, o[5]) * f6 + at(D2, o[6]) * f7 + at(D2, o[7]) * f8);
    // dis2 / (float) N^3 (main.c:1503): for a power-of-two N the reciprocal is exact and the
    // product equals the quotient; otherwise the correctly rounded ddiv
    dis2[a] = cc.n3pow2 ? dis2[a] * cc.rn3 : ddiv(dis2[a], cc.n3f, cc.rn3);
  }
  float V[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    P[a] = (float)((double)P[a] + (dis[a] - 3. / 7. * dis2[a]));
    V[a] = (float)(dis[a] * vp1 - 3. / 7. * dis2[a] * vp2);
    P[a] = (float)fmod_box(fmod_box((double)P[a], box) + box, box);
  }
  // one 12-byte store per array: a wavefront writes 768 contiguous bytes
  reinterpret_cast<float3*>(pos)[p] = make_float3(P[0], P[1], P[2]);
  reinterpret_cast<float3*>(vel)[p] = make_float3(V[0], V[1], V[2]);
}

// Block (b, t) covers 256 consecutive glass particles b of tile t: with the lattice glass, four
// z-rows at one x, so blocks b and b + 1 share a y-row of corners and b and b + 16 an x-plane.
// Workgroups are dispatched round-robin over the 8 XCDs (linear id w on XCD w % 8); within each
// run of 16 linear ids the blocks are permuted so that XCD k takes the adjacent blocks 2k and
// 2k + 1, which share a row of corners in its L2 (in the natural order every y-neighbour sits on
// another XCD), while blocks 16 apart stay on one XCD as before.  Giving each XCD a contiguous
// range of blocks instead read more (at 512^3 12.2 against 8.6 spectra, 2.17 against 2.01 ms):
// the XCDs then work on distant parts of the particle array at once.
template <typename T>
__global__ void __launch_bounds__(256)
    k_cic(const cx_t<T>* __restrict__ disp, const cx_t<T>* __restrict__ disp2, Geom g,
          const float* __restrict__ glass, int nglass, CicConst cc, float* __restrict__ pos, float* __restrict__ vel) {
  unsigned w = blockIdx.x + blockIdx.y * gridDim.x;
  if ((gridDim.x * gridDim.y) % 16 == 0) w = (w & ~15u) | ((w & 7u) << 1) | ((w >> 3) & 1u);
  const int n = (int)(w % gridDim.x) * blockDim.x + threadIdx.x;  // glass particle
  if (n >= nglass) return;
  cic_particle<T>(n, (int)(w / gridDim.x), disp, disp2, g, glass, nglass, cc, pos, vel);
}

template <typename T>
int Pipeline<T>::cic(const cx_t<T>* disp3, const cx_t<T>* disp2_3, const Geom& g, const float* glass,
                     int nglass, double gbox, int tile, double box, double vp1, double vp2, float* pos,
                     float* vel, long long npart, hipStream_t st) {
  const long long ntile = (long long)tile * tile * tile;
  if (ntile * nglass != npart || ntile > 65535) return AAR_ERR_UNSUPPORTED;
  CicConst cc;
  cc.gbox = gbox, cc.rgbox = 1.0 / gbox, cc.box = box, cc.rbox = 1.0 / box, cc.bt = box / tile;
  cc.vp1 = vp1, cc.vp2 = vp2, cc.tile = tile;
  cc.n3f = (double)(float)((long long)g.N * g.N * g.N);
  cc.rn3 = 1.0 / cc.n3f;
  cc.n3pow2 = g.pow2;
  hipLaunchKernelGGL((k_cic<T>), dim3((unsigned)((nglass + 255) / 256), (unsigned)ntile), dim3(256), 0, st,
                     disp3, disp2_3, g, glass, nglass, cc, pos, vel);
  return hipGetLastError() == hipSuccess ? AAR_OK : AAR_ERR_HIP;
}

// Plain streaming copy of n 16-byte words: the reference kernel for the
// measured HBM copy peak and for scaling the FETCH_SIZE / WRITE_SIZE counters.
// One word per lane: on MI355X this shape copies faster than several words
// per lane (6.2 against 5.5 TB/s over 4.3 GB).
__global__ void __launch_bounds__(256) k_copy16(const float4* __restrict__ src, float4* __restrict__ dst,
                                                long long n) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

int aar_copy16(const void* src, void* dst, long long bytes, hipStream_t st) {
  if (bytes % 16) return AAR_ERR_BAD_ARG;
  const long long n = bytes / 16;
  const long long blocks = (n + 255) / 256;
  if (blocks > 0x7fffffffLL) return AAR_ERR_UNSUPPORTED;
  hipLaunchKernelGGL(k_copy16, dim3((unsigned)blocks), dim3(256), 0, st, (const float4*)src, (float4*)dst, n);
  return hipGetLastError() == hipSuccess ? AAR_OK : AAR_ERR_HIP;
}

// The x / y passes' memory access without the transform: each workgroup reads one 128-byte
// column tile (16 adjacent kz columns of FP32 values, or 8 of FP64, at every position along the
// axis, the addresses of tile shape 0) into registers and writes it back in place.  Its rate
// bounds what a pass with that access pattern can reach (the reference for the x / y passes'
// roofline at 1024^3, where the tile no longer stays in the last-level cache).
template <typename T, int AXIS>
__global__ void __launch_bounds__(1024) k_xy_pattern(cx_t<T>* __restrict__ f, int N, int H) {
  constexpr int C = sizeof(T) == 4 ? 16 : 8, E = 16;
  const int t = blockIdx.x, per = H / C, a = t / per, c0 = (t % per) * C;
  const int col = threadIdx.x % C, p0 = threadIdx.x / C, rows = N / E;
  cx_t<T>* fb = f + (long long)blockIdx.y * ((long long)N * N * (H + 1));
  cx_t<T> v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int pos = p0 + e * rows;
    v[e] = fb[(AXIS == 0 ? row_of(N, pos, a) : row_of(N, a, pos)) * H + c0 + col];
  }
  __syncthreads();
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int pos = p0 + e * rows;
    fb[(AXIS == 0 ? row_of(N, pos, a) : row_of(N, a, pos)) * H + c0 + col] = v[e];
  }
}

template <typename T> int aar_xy_pattern(cx_t<T>* f, int nf, const Geom& g, int axis, hipStream_t st) {
  constexpr int C = sizeof(T) == 4 ? 16 : 8;
  if (!g.pow2 || g.N < 2 * C || g.N / 16 * C > 1024) return AAR_ERR_UNSUPPORTED;
  const dim3 grid(g.N * (g.H / C), nf), block(C * g.N / 16);
  if (axis == 0) hipLaunchKernelGGL((k_xy_pattern<T, 0>), grid, block, 0, st, f, g.N, g.H);
  else hipLaunchKernelGGL((k_xy_pattern<T, 1>), grid, block, 0, st, f, g.N, g.H);
  return hipGetLastError() == hipSuccess ? AAR_OK : AAR_ERR_HIP;
}
template int aar_xy_pattern<float>(float2*, int, const Geom&, int, hipStream_t);
template int aar_xy_pattern<double>(double2*, int, const Geom&, int, hipStream_t);

__global__ void k_zero(double* p, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

int aar_zero_fill(double* p, long long n, hipStream_t st) {
  hipLaunchKernelGGL(k_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n);
  return hipGetLastError() == hipSuccess ? AAR_OK : AAR_ERR_HIP;
}

// n values of a fixed hash pattern in [-1/2, 1/2): timing data for the autotuner (FFT passes
// over non-zero data, as in a job)
template <typename T> __global__ void k_noise(T* p, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned h = (unsigned)(i * 2654435761ull) ^ (unsigned)(i >> 29);
  p[i] = (T)((h >> 8) & 0xffffu) * (T)(1.0 / 65536) - (T)0.5;
}

template <typename T> int aar_noise_fill(T* p, long long n, hipStream_t st) {
  hipLaunchKernelGGL((k_noise<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n);
  return hipGetLastError() == hipSuccess ? AAR_OK : AAR_ERR_HIP;
}
template int aar_noise_fill<float>(float*, long long, hipStream_t);
template int aar_noise_fill<double>(double*, long long, hipStream_t);

__global__ void k_ranlxd_draws(uint32_t seed, int n, double* out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  Ranlxd r;
  ranlxd_seed(r, seed);
  for (int b = 0; b < n; b += 12) {
    ranlxd_update_dev(r);
#pragma unroll
    for (int q = 0; q < 12; ++q)
      if (b + q < n) out[b + q] = (double)r.s[q] * kTwoM48;
  }
}

int aar_ranlxd_draws_dev(uint32_t seed, int n, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_ranlxd_draws, dim3(1), dim3(64), 0, st, seed, n, out);
  return hipGetLastError() == hipSuccess ? AAR_OK : AAR_ERR_HIP;
}

template struct Pipeline<float>;
template struct Pipeline<double>;
