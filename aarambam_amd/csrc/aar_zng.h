// Fused z pass of a product sum, shared by the kernels built with the library and the ones
// compiled at run time for a job's own product schedule (aar_jit.cpp).
#pragma once
#include "aar_fft.h"

// r2c post-processing of C transformed rows held in a row-major tile:
// X_k = E_k + e^{-2 pi i k/N} O_k from Z_k and Z_{H-k} (E = (Z_k + Z*_{H-k})/2,
// O = (Z_k - Z*_{H-k})/(2i)), stored along k (coalesced).  Unit k in [1, H/2)
// writes bins k and H - k; unit 0 writes bins 0 and H/2 and the Nyquist plane
// entry.  With wavefront-local tiles each wavefront handles its own rows.
template <typename T, int H, int C, int NT, bool WL>
__device__ __forceinline__ void r2c_post(const cx_t<T>* tile, const cx_t<T>* tw, cx_t<T>* f, long long r0,
                                         const Geom& g) {
  using C2 = cx_t<T>;
  using TI = TileIdx<T, H, C, true, C * H / NT>;
  C2* rows = f + r0 * H;
  constexpr int U = H / 2;
  constexpr int GT = WL ? 64 : NT, GC = WL ? C * 64 / NT : C;
  const int gt = WL ? (int)(threadIdx.x & 63) : (int)threadIdx.x;
  const int gc0 = WL ? (int)(threadIdx.x >> 6) * GC : 0;
  auto bin = [&](int k, const C2 A, const C2 B) {  // X_k from Z_k = A and Z_{H-k} = B
    const C2 Ev = cscale(cadd(A, cconj(B)), (T)0.5);
    const C2 dd = csub(A, cconj(B));
    const C2 O = cmake<C2>(dd.y * (T)0.5, -dd.x * (T)0.5);
    return cadd(Ev, cmul(O, tw[k]));
  };
  for (int e = gt; e < GC * U; e += GT) {
    const int c = gc0 + e / U, k = e % U;
    C2* out = rows + (long long)c * H;
    if (k == 0) {
      const C2 z = tile[TI::at(0, c)];
      out[0] = cmake<C2>(z.x + z.y, (T)0);
      f[g.nq + r0 + c] = cmake<C2>(z.x - z.y, (T)0);
      const C2 m = tile[TI::at(U, c)];
      out[U] = bin(U, m, m);
    } else {
      const int k2 = H - k;
      const C2 A = tile[TI::at(k, c)], B = tile[TI::at(k2, c)];
      out[k] = bin(k, A, B);
      out[k2] = bin(k2, B, A);
    }
  }
}

// The paired form applies to FP32 at H = 128 and 256 (z_ng_body_paired, below).
template <typename T, int H, int FIX> struct ZPair {
  static constexpr bool value = FIX == 0 && sizeof(T) == 4 && (H == 128 || H == 256);
  static constexpr int N = 2 * H, EF = 8, C = 4, NT = C * N / EF;
  static constexpr bool WL = wave_local(N, EF) && NT % 64 == 0;
};

// For C grid rows at a time: half-complex -> real along z for each of the M
// mode fields (kept in registers), the products h_K = sum c f_i f_j of the
// regrouped NG sum, and real -> half-complex along z of each h_K, written
// over the first K field slots.  Replaces z-C2R of M fields + the product
// kernel + z-R2C of K fields ((M + K) spectra of HBM traffic instead of
// 3(M + K)).
template <typename T, int H, int MMAX, int FIX = 0> struct ZNGCfg {
  // complex elements per thread: 8 keeps the 1024-point rows in one
  // wavefront each when at most 6 mode fields sit in registers.  The 2LPT
  // source (FIX) holds at most three fields at once and keeps 4: its 118
  // VGPRs give four waves per SIMD where 8 elements took 266 (one wave), and
  // the kernel runs 13 % faster at 1024^3 for its workgroup barriers
  static constexpr int E = sizeof(T) == 4 ? (H == 512 && MMAX <= 6 && !FIX ? 8 : 4) : 2;
  static constexpr int C = sizeof(T) == 4 && (H < 512 || E == 8) ? 4 : 2;  // rows per block
  static constexpr int NT = C * H / E;
  // every row's exchanges inside one wavefront: no workgroup barriers
  static constexpr bool WL = wave_local(H, E) && NT % 64 == 0;
  // the paired form holds two fields' raw inputs and their prefetch: its allocation is held to
  // 256 registers (two waves per SIMD, one spilled value) where it would take 262 (one wave)
  static constexpr int WPE = ZPair<T, H, FIX>::value ? 2 : 1;
};

// FIX = 0: the NG sum, M mode fields and K products, those of output kk formed by
// prod.operator()<E>(kk, get, hv) from get(m, v) (the values of mode slot m at the thread's E
// slots) into hv;
// FIX = 1: the 2LPT source, src = f00 (f11 + f22) + f11 f22 - f01^2 - f02^2 - f12^2
// (main.c:1286-1290) of the six gradient fields 00, 01, 02, 11, 12, 22 in slots 0-5,
// one product written over slot 0.  With FIX the mode loop is unrolled, so every
// register index is a constant (no indexed register moves) and the product is
// the reference's expression in its order.
template <typename T, int H, int MMAX, int FIX, class Prod>
__device__ __forceinline__ void z_ng_body(cx_t<T>* base, int M, int K, const Geom& g,
                                          const cx_t<T>* __restrict__ twg, const Prod& prod) {
  using C2 = cx_t<T>;
  using Cfg = ZNGCfg<T, H, MMAX, FIX>;
  constexpr int C = Cfg::C, NT = Cfg::NT, E = Cfg::E;
  constexpr bool WL = Cfg::WL;
  constexpr int N = 2 * H;
  static_assert(!FIX || MMAX == 6, "the 2LPT source has six fields");
  using TI = TileIdx<T, H, C, true, C * H / NT>;
  __shared__ C2 tile[TI::SIZE];
  __shared__ C2 tw[N];
  load_twiddles(tw, twg + N, N, NT);
  __syncthreads();
  const long long r0 = (long long)blockIdx.x * C;
  // Real-space values of the mode fields: (x[2 pos], x[2 pos + 1]) of slot s
  // in fx<s>[m], fy<s>[m].  One small array per slot and component, each
  // indexed by the uniform mode number, stays in VGPRs (indirect addressing).
  static_assert(E == 2 || E == 4 || E == 8, "two, four or eight slots per thread");
  T fx0[MMAX], fx1[MMAX], fx2[MMAX], fx3[MMAX], fy0[MMAX], fy1[MMAX], fy2[MMAX], fy3[MMAX];
  T fx4[MMAX], fx5[MMAX], fx6[MMAX], fx7[MMAX], fy4[MMAX], fy5[MMAX], fy6[MMAX], fy7[MMAX];
  // Raw inputs of the first FFT stage (X[k], X[H-k] per slot, X[H] for k = 0)
  // are prefetched one field ahead, so HBM loads overlap the previous field's
  // transform.
  C2 ra[E], rb[E], na[E], nb[E];
  auto fetch = [&](int m, C2* a, C2* b) {
    const C2* fm = base + (long long)m * g.fs;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      int pos, col;
      first_slot<T, H, C, NT, true>(e, pos, col);
      const C2* row = fm + (r0 + col) * H;
      a[e] = row[pos];
      b[e] = pos == 0 ? fm[g.nq + r0 + col] : row[H - pos];
    }
  };
  auto c2r_mode = [&](int m, int next) {  // next: the slot to prefetch, or -1
    // keeps the scheduler from hoisting later fields' loads (unrolled mode loop) above this one
    __builtin_amdgcn_sched_barrier(0);
    if (next >= 0) fetch(next, na, nb);
    auto load = [&](int k, int, int e) {
      const C2 A = ra[e], B = rb[e];
      if (k == 0) return cmake<C2>(A.x + B.x, A.x - B.x);  // B = X[H] here
      C2 w = tw[k];
      w.y = -w.y;
      const C2 s = cadd(A, cconj(B)), d = csub(A, cconj(B));
      const C2 dw = cmul(d, w);
      return cmake<C2>(s.x - dw.y, s.y + dw.x);
    };
    auto store = [&](int, int, C2 v, int e) {
      if (e == 0) { fx0[m] = v.x; fy0[m] = v.y; }
      if (e == 1) { fx1[m] = v.x; fy1[m] = v.y; }
      if constexpr (E >= 4) {
        if (e == 2) { fx2[m] = v.x; fy2[m] = v.y; }
        if (e == 3) { fx3[m] = v.x; fy3[m] = v.y; }
      }
      if constexpr (E == 8) {
        if (e == 4) { fx4[m] = v.x; fy4[m] = v.y; }
        if (e == 5) { fx5[m] = v.x; fy5[m] = v.y; }
        if (e == 6) { fx6[m] = v.x; fy6[m] = v.y; }
        if (e == 7) { fx7[m] = v.x; fy7[m] = v.y; }
      }
    };
    fft_tile<T, H, C, NT, +1, 2, false, true, WL>(tile, tw, load, store);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      ra[e] = na[e];
      rb[e] = nb[e];
    }
  };
  // every read of a mode slot is a run of register reads with one uniform index
  auto get = [&](int m, C2* v) {
    v[0] = cmake<C2>(fx0[m], fy0[m]);
    v[1] = cmake<C2>(fx1[m], fy1[m]);
    if constexpr (E >= 4) {
      v[2] = cmake<C2>(fx2[m], fy2[m]);
      v[3] = cmake<C2>(fx3[m], fy3[m]);
    }
    if constexpr (E == 8) {
      v[4] = cmake<C2>(fx4[m], fy4[m]);
      v[5] = cmake<C2>(fx5[m], fy5[m]);
      v[6] = cmake<C2>(fx6[m], fy6[m]);
      v[7] = cmake<C2>(fx7[m], fy7[m]);
    }
  };
  // z R2C of one product (values hv at the thread's last-stage slots) into slot kk
  auto r2c_product = [&](int kk, const C2* hv) {
    tile_sync<WL>();  // the tile is free
#pragma unroll
    for (int e = 0; e < E; ++e) {
      int pos, col;
      last_slot<T, H, C, NT, true>(e, pos, col);
      tile[TI::at(pos, col)] = hv[e];
    }
    auto load = [&](int pos, int c) { return tile[TI::at(pos, c)]; };
    auto store = [&](int pos, int c, C2 v) { tile[TI::at(pos, c)] = v; };
    fft_tile<T, H, C, NT, -1, 2, true, true, WL>(tile, tw, load, store);
    tile_sync<WL>();
    r2c_post<T, H, C, NT, WL>(tile, tw, base + (long long)kk * g.fs, r0, g);
  };
  if constexpr (FIX == 1) {
    // fields in the order the expression consumes them (00, 11, 22, then 01, 02, 12), so at
    // most three are held at once and each square is subtracted as soon as its field is ready
    C2 a00[E], a11[E], a22[E], hv[E];
    fetch(0, ra, rb);
    c2r_mode(0, 3);
    c2r_mode(3, 5);
    c2r_mode(5, 1);
    get(0, a00);
    get(3, a11);
    get(5, a22);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      hv[e].x = a00[e].x * (a11[e].x + a22[e].x) + a11[e].x * a22[e].x;
      hv[e].y = a00[e].y * (a11[e].y + a22[e].y) + a11[e].y * a22[e].y;
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {  // slots 1, 2, 4: f01, f02, f12
      const int m = q == 2 ? 4 : q + 1;
      c2r_mode(m, q == 0 ? 2 : q == 1 ? 4 : -1);
      get(m, a00);
#pragma unroll
      for (int e = 0; e < E; ++e) {
        hv[e].x = hv[e].x - a00[e].x * a00[e].x;
        hv[e].y = hv[e].y - a00[e].y * a00[e].y;
      }
    }
    r2c_product(0, hv);
    return;
  } else {
    fetch(0, ra, rb);
    for (int m = 0; m < M; ++m) c2r_mode(m, m + 1 < M ? m + 1 : -1);
  }
  for (int kk = 0; kk < K; ++kk) {
    C2 hv[E];
#pragma unroll
    for (int e = 0; e < E; ++e) hv[e] = cmake<C2>(0, 0);
    prod.template operator()<E>(kk, get, hv);
    r2c_product(kk, hv);
  }
}


// Paired form of the NG sum's fused z pass (FP32, H = 128 or 256, the 256^3 and 512^3 meshes).
// Two mode fields a, b go through one N-point complex inverse transform of Y = A + i B, with A
// and B the Hermitian half spectra (Y_p = A_p + i B_p for p <= H, conj(A_{N-p}) + i conj(B_{N-p})
// above; the imaginary parts of X_0 and X_H are ignored as FFTW's half-complex convention
// does): the real and imaginary parts of the result are a and b.  Two products h1, h2 go
// through one forward transform of h1 + i h2, split as H1_k = (Z_k + Z*_{N-k}) / 2 and
// H2_k = (Z_k - Z*_{N-k}) / (2 i).  Per field this drops the half-length transform's pre- and
// post-twiddles and one LDS exchange; the launch shape is the unpaired one's (C = 4 rows,
// C N / 8 threads: eight elements per thread, one wavefront per 512-point row).

template <typename T, int H, int MMAX, class Prod>
__device__ __forceinline__ void z_ng_body_paired(cx_t<T>* base, int M, int K, const Geom& g,
                                                 const cx_t<T>* __restrict__ twg, const Prod& prod) {
  using C2 = cx_t<T>;
  using Cfg = ZPair<T, H, 0>;
  constexpr int N = Cfg::N, EF = Cfg::EF, C = Cfg::C, NT = Cfg::NT;
  constexpr bool WL = Cfg::WL;
  constexpr int E = EF / 2;  // the products see the eight real samples as four pairs
  static_assert(NT == ZNGCfg<T, H, MMAX, 0>::NT && C == ZNGCfg<T, H, MMAX, 0>::C, "launch shape");
  using TI = TileIdx<T, N, C, true, EF>;
  __shared__ C2 tile[TI::SIZE];
  __shared__ C2 tw[N];
  load_twiddles(tw, twg + N, N, NT);
  __syncthreads();
  const long long r0 = (long long)blockIdx.x * C;
  // real samples of the mode fields at the thread's EF last-stage positions: sample e of slot m
  // in fx<e/2>[m] (e even) or fy<e/2>[m] (e odd), indexed by the uniform mode number
  T fx0[MMAX], fx1[MMAX], fx2[MMAX], fx3[MMAX], fy0[MMAX], fy1[MMAX], fy2[MMAX], fy3[MMAX];
  auto put = [&](int m, int e, T v) {
    if (e == 0) fx0[m] = v;
    if (e == 1) fy0[m] = v;
    if (e == 2) fx1[m] = v;
    if (e == 3) fy1[m] = v;
    if (e == 4) fx2[m] = v;
    if (e == 5) fy2[m] = v;
    if (e == 6) fx3[m] = v;
    if (e == 7) fy3[m] = v;
  };
  auto get = [&](int m, C2* v) {
    v[0] = cmake<C2>(fx0[m], fy0[m]);
    v[1] = cmake<C2>(fx1[m], fy1[m]);
    v[2] = cmake<C2>(fx2[m], fy2[m]);
    v[3] = cmake<C2>(fx3[m], fy3[m]);
  };
  // X_p of field m at the thread's first-stage positions p (X_{N-p} above H, X_H from the
  // Nyquist plane), prefetched one pair ahead
  C2 ra[EF], rb[EF], na[EF], nb[EF];
  auto fetch = [&](int m, C2* out) {
    const C2* fm = base + (long long)m * g.fs;
#pragma unroll
    for (int e = 0; e < EF; ++e) {
      int pos, col;
      first_slot<T, N, C, NT, true>(e, pos, col);
      const C2* row = fm + (r0 + col) * H;
      out[e] = pos < H ? row[pos] : (pos == H ? fm[g.nq + r0 + col] : row[N - pos]);
    }
  };
  auto zero = [&](C2* out) {
#pragma unroll
    for (int e = 0; e < EF; ++e) out[e] = cmake<C2>(0, 0);
  };
  // fields m and m + 1 (m + 1 == M: a zero field) -> slots m, m + 1; prefetch the pair at next
  auto c2r_pair = [&](int m, int next) {
    __builtin_amdgcn_sched_barrier(0);
    if (next >= 0) {
      fetch(next, na);
      if (next + 1 < M) fetch(next + 1, nb);
      else zero(nb);
    }
    auto load = [&](int p, int, int e) {
      const C2 A = ra[e], B = rb[e];
      if (p == 0 || p == H) return cmake<C2>(A.x, B.x);
      if (p < H) return cmake<C2>(A.x - B.y, A.y + B.x);
      return cmake<C2>(A.x + B.y, B.x - A.y);
    };
    auto store = [&](int, int, C2 v, int e) {
      put(m, e, v.x);
      if (m + 1 < MMAX) put(m + 1, e, v.y);
    };
    fft_tile<T, N, C, NT, +1, 1, false, true, WL>(tile, tw, load, store);
#pragma unroll
    for (int e = 0; e < EF; ++e) {
      ra[e] = na[e];
      rb[e] = nb[e];
    }
  };
  // products kk and kk + 1 (kk + 1 == K: none) -> slots kk, kk + 1
  auto r2c_pair = [&](int kk, const C2* h1, const C2* h2) {
    tile_sync<WL>();  // the tile is free
#pragma unroll
    for (int e = 0; e < EF; ++e) {
      int pos, col;
      last_slot<T, N, C, NT, true>(e, pos, col);
      const T a = (e & 1) ? h1[e / 2].y : h1[e / 2].x, b = (e & 1) ? h2[e / 2].y : h2[e / 2].x;
      tile[TI::at(pos, col)] = cmake<C2>(a, b);
    }
    auto load = [&](int pos, int c) { return tile[TI::at(pos, c)]; };
    auto store = [&](int pos, int c, C2 v) { tile[TI::at(pos, c)] = v; };
    fft_tile<T, N, C, NT, -1, 1, true, true, WL>(tile, tw, load, store);
    tile_sync<WL>();
    // bins k in [0, H] of the thread's rows (with wavefront-local rows, of its wavefront's rows)
    constexpr int GT = WL ? 64 : NT, GC = WL ? C * 64 / NT : C;
    const int gt = WL ? (int)(threadIdx.x & 63) : (int)threadIdx.x;
    const int gc0 = WL ? (int)(threadIdx.x >> 6) * GC : 0;
    C2* f1 = base + (long long)kk * g.fs;
    C2* f2 = base + (long long)(kk + 1) * g.fs;
    const bool two = kk + 1 < K;
    for (int q = gt; q < GC * (H + 1); q += GT) {
      const int c = gc0 + q / (H + 1), k = q % (H + 1);
      const C2 Z = tile[TI::at(k, c)], W = tile[TI::at(k == 0 ? 0 : N - k, c)];
      const C2 s = cmake<C2>((Z.x + W.x) * (T)0.5, (Z.y - W.y) * (T)0.5);  // (Z + W*) / 2
      const C2 d = cmake<C2>((Z.y + W.y) * (T)0.5, (W.x - Z.x) * (T)0.5);  // (Z - W*) / (2 i)
      const long long o = k < H ? (r0 + c) * H + k : g.nq + r0 + c;
      f1[o] = s;
      if (two) f2[o] = d;
    }
  };
  fetch(0, ra);
  if (1 < M) fetch(1, rb);
  else zero(rb);
  for (int m = 0; m < M; m += 2) c2r_pair(m, m + 2 < M ? m + 2 : -1);
  for (int kk = 0; kk < K; kk += 2) {
    C2 h1[E], h2[E];
#pragma unroll
    for (int e = 0; e < E; ++e) h1[e] = h2[e] = cmake<C2>(0, 0);
    prod.template operator()<E>(kk, get, h1);
    if (kk + 1 < K) prod.template operator()<E>(kk + 1, get, h2);
    r2c_pair(kk, h1, h2);
  }
}

// The fused z pass of a product sum: the paired form where it applies, else z_ng_body.
template <typename T, int H, int MMAX, int FIX, class Prod>
__device__ __forceinline__ void z_ng_run(cx_t<T>* base, int M, int K, const Geom& g, const cx_t<T>* __restrict__ twg,
                                         const Prod& prod) {
  if constexpr (ZPair<T, H, FIX>::value) z_ng_body_paired<T, H, MMAX>(base, M, K, g, twg, prod);
  else z_ng_body<T, H, MMAX, FIX>(base, M, K, g, twg, prod);
}
