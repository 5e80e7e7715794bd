// Batched Stockham FFT on an LDS tile for gfx950.
//
// A tile holds C independent transforms ("columns") of length L, element
// (pos, col) at tile[pos*CP + col] (CP = C + 1, padded against bank
// conflicts).  NT threads cooperate; thread t owns column t % C and
// E = C*L/NT elements per stage.  Each stage of radix R performs, for every
// butterfly index i in [0, L/R) with k = i mod P (P = product of the earlier
// radices):  v_t = x[i + t L/R] * w^(t k L/(P R)),  y = DFT_R(v),
// x'[(i-k) R + k + u P] = y_u   (Stockham autosort, no bit reversal).
// The first stage reads through a caller-supplied loader (global memory or
// LDS) and the last stage writes through a caller-supplied storer, so loads
// and stores to HBM are coalesced across the C adjacent columns and the
// intermediate stages stay in LDS.
#pragma once
#include "aar_device.h"

// Loaders may take (pos, col) or (pos, col, e), e = the thread's element slot; storers
// (pos, col, v) or (pos, col, v, e), e = the thread's output slot.  The form with the slot is
// chosen when it exists (overload on int, the exact match, over long).
template <class F> __device__ __forceinline__ auto call_load_(F& f, int pos, int col, int e, int) -> decltype(f(pos, col, e)) {
  return f(pos, col, e);
}
template <class F> __device__ __forceinline__ auto call_load_(F& f, int pos, int col, int, long) -> decltype(f(pos, col)) {
  return f(pos, col);
}
template <class F> __device__ __forceinline__ auto call_load(F& f, int pos, int col, int e) {
  return call_load_(f, pos, col, e, 0);
}
template <class F, class V>
__device__ __forceinline__ auto call_store_(F& f, int pos, int col, V v, int e, int) -> decltype(f(pos, col, v, e), void()) {
  f(pos, col, v, e);
}
template <class F, class V>
__device__ __forceinline__ auto call_store_(F& f, int pos, int col, V v, int, long) -> decltype(f(pos, col, v), void()) {
  f(pos, col, v);
}
template <class F, class V> __device__ __forceinline__ void call_store(F& f, int pos, int col, V v, int e) {
  call_store_(f, pos, col, v, e, 0);
}

__host__ __device__ constexpr int ilog2c(int v) { return v <= 1 ? 0 : 1 + ilog2c(v / 2); }

// Radix plans of ceil(log2 L / log2 Rmax) stages (the fewest LDS exchanges):
//   PL = 0: the bits spread evenly, larger radices first (512 = 8 8 8, 1024 = 16 8 8);
//   PL = 1: radix Rmax for every stage but the last, which takes the rest (16 16 2, 16 16 4);
//   PL = 2: the rest first, then radix Rmax (2 16 16, 4 16 16).
// The plans differ in rounding only; the autotuner picks one per precision and mesh size for
// the x / y passes.
__host__ __device__ constexpr int plan_nstages(int L, int Rmax) {
  return (ilog2c(L) + ilog2c(Rmax) - 1) / ilog2c(Rmax);
}
__host__ __device__ constexpr int plan_radix(int L, int Rmax, int s, int PL = 0) {
  return PL == 1 ? (s < plan_nstages(L, Rmax) - 1 ? Rmax : 1 << (ilog2c(L) - ilog2c(Rmax) * (plan_nstages(L, Rmax) - 1)))
       : PL == 2 ? (s > 0 ? Rmax : 1 << (ilog2c(L) - ilog2c(Rmax) * (plan_nstages(L, Rmax) - 1)))
                 : 1 << (ilog2c(L) / plan_nstages(L, Rmax) + (s < ilog2c(L) % plan_nstages(L, Rmax) ? 1 : 0));
}
__host__ __device__ constexpr int plan_prefix(int L, int Rmax, int s, int PL = 0) {
  return s == 0 ? 1 : plan_prefix(L, Rmax, s - 1, PL) * plan_radix(L, Rmax, s - 1, PL);
}

// LDS row pitch of a tile: one padding column when the tile plus its twiddle
// table stay below 152 KiB of the 160 KiB of LDS of a gfx950 CU.
template <typename T> __host__ __device__ constexpr int tile_cp(int L, int C) {
  return ((long)L * (C + 1) + L) * (long)sizeof(cx_t<T>) <= 152 * 1024 ? C + 1 : C;
}

// ---------------------------------------------------------------- DFT kernels
template <int DIR, typename C> __device__ __forceinline__ void dft2(C& a, C& b) {
  C t = a;
  a = cadd(t, b);
  b = csub(t, b);
}

template <int DIR, typename C> __device__ __forceinline__ void dft4(C& x0, C& x1, C& x2, C& x3) {
  C a = cadd(x0, x2), b = csub(x0, x2), c = cadd(x1, x3), d = cmul_i<DIR>(csub(x1, x3));
  x0 = cadd(a, c);
  x2 = csub(a, c);
  x1 = cadd(b, d);
  x3 = csub(b, d);
}

template <int DIR, typename C> __device__ __forceinline__ void dft4v(C* v) { dft4<DIR>(v[0], v[1], v[2], v[3]); }

template <int DIR, typename C> __device__ __forceinline__ void dft8v(C* v) {
  using T = decltype(C::x);
  const T h = (T)0.70710678118654752440;
  C e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  C o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4<DIR>(e0, e1, e2, e3);
  dft4<DIR>(o0, o1, o2, o3);
  // twiddles W8^u, W8 = exp(DIR * 2 pi i / 8)
  C t1 = cmake<C>(h * (o1.x - (T)DIR * o1.y), h * (o1.y + (T)DIR * o1.x));
  C t2 = cmul_i<DIR>(o2);
  C t3 = cmake<C>(h * (-o3.x - (T)DIR * o3.y), h * (-o3.y + (T)DIR * o3.x));
  v[0] = cadd(e0, o0);
  v[4] = csub(e0, o0);
  v[1] = cadd(e1, t1);
  v[5] = csub(e1, t1);
  v[2] = cadd(e2, t2);
  v[6] = csub(e2, t2);
  v[3] = cadd(e3, t3);
  v[7] = csub(e3, t3);
}

template <int DIR, typename C> __device__ __forceinline__ void dft16v(C* v) {
  using T = decltype(C::x);
  // n = 4 n1 + n2, u = u1 + 4 u2
  C a[4][4];
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) {
    a[n2][0] = v[n2];
    a[n2][1] = v[4 + n2];
    a[n2][2] = v[8 + n2];
    a[n2][3] = v[12 + n2];
    dft4<DIR>(a[n2][0], a[n2][1], a[n2][2], a[n2][3]);
  }
  // twiddle W16^(n2 u1)
  const T c1 = (T)0.92387953251128675613, s1 = (T)0.38268343236508977173;
  const T h = (T)0.70710678118654752440;
#pragma unroll
  for (int n2 = 1; n2 < 4; ++n2) {
#pragma unroll
    for (int u1 = 1; u1 < 4; ++u1) {
      const int m = n2 * u1;  // 1..9, angle DIR * 2 pi m / 16
      T cr, ci;
      switch (m) {
        case 1: cr = c1; ci = s1; break;
        case 2: cr = h; ci = h; break;
        case 3: cr = s1; ci = c1; break;
        case 4: cr = 0; ci = 1; break;
        case 6: cr = -h; ci = h; break;
        case 9: cr = -c1; ci = -s1; break;
        default: cr = 1; ci = 0; break;
      }
      ci *= (T)DIR;
      C x = a[n2][u1];
      a[n2][u1] = cmake<C>(x.x * cr - x.y * ci, x.x * ci + x.y * cr);
    }
  }
#pragma unroll
  for (int u1 = 0; u1 < 4; ++u1) {
    C b0 = a[0][u1], b1 = a[1][u1], b2 = a[2][u1], b3 = a[3][u1];
    dft4<DIR>(b0, b1, b2, b3);
    v[u1] = b0;
    v[u1 + 4] = b1;
    v[u1 + 8] = b2;
    v[u1 + 12] = b3;
  }
}

template <int R, int DIR, typename C> __device__ __forceinline__ void dftR(C* v) {
  if constexpr (R == 2) dft2<DIR>(v[0], v[1]);
  else if constexpr (R == 4) dft4v<DIR>(v);
  else if constexpr (R == 8) dft8v<DIR>(v);
  else dft16v<DIR>(v);
}

// ---------------------------------------------------------------- tile FFT
// Two tile layouts:
//  column-major (RM = false): element (pos, col) at pos*CP + col, thread t
//    serves column t % C, so a wavefront touches 16 adjacent columns at a few
//    positions -- coalesced when columns are adjacent in memory (x, y passes);
//  row-major (RM = true): element at col*LP + pos + pad(pos), thread t serves
//    butterfly t % (L/R) of column t / (L/R), so a wavefront runs along
//    positions -- coalesced when each column is a contiguous row (z passes).
// The LDS of gfx950 serves a 64-bit access in passes of 32 lanes over 32
// eight-byte banks.  pad() is a sum of shifted position bits: the butterfly
// reads (i + t L/R) and Stockham writes ((i - k) R + k + u P) split into a
// per-lane part and a constant whose bits do not overlap, so pad() of the
// sum is pad(lane part) + pad(constant) and the constant folds into the
// instruction's offset.  For FP32 rows with E = 4 or 8 elements per thread,
// 2 (pos >> 5) + (pos >> 6) or 2 (pos >> 5) + (pos >> 7) keeps most of the
// writes' lanes on distinct banks, where pos >> 4 left every write and most
// reads of these plans two-way conflicted; other tiles keep pos >> 4.
template <typename T, int L, int C, bool RM, int E = 0> struct TileIdx {
  static constexpr int PADK = (RM && sizeof(T) == 4 && (E == 4 || E == 8)) ? E : 0;
  __host__ __device__ static constexpr int pad(int pos) {
    if constexpr (PADK == 4) return 2 * (pos >> 5) + (pos >> 6);
    else if constexpr (PADK == 8) return 2 * (pos >> 5) + (pos >> 7);
    else return pos >> 4;
  }
  static constexpr int CP = tile_cp<T>(L, C);
  static constexpr int LP = L + pad(L - 1) + 1 + (PADK && L == 64 ? 2 : 0) + (PADK && L == 32 ? 4 : 0);
  static constexpr int SIZE = RM ? C * LP : L * CP;
  __device__ static __forceinline__ int at(int pos, int col) {
    if constexpr (RM) return col * LP + pos + pad(pos);
    else return pos * CP + col;
  }
};

// Row-major tiles in which every stage has radix E (one butterfly per thread,
// thread t on column t / (L/E)) and L/E divides 64: each column belongs to one
// wavefront at every stage, so the exchanges between stages stay inside a
// wavefront, whose LDS operations complete in issue order, and a
// wavefront-scope fence replaces the workgroup barrier.
__host__ __device__ constexpr bool wave_local(int L, int E) {
  if (E > 16) return false;
  for (int s = 0; s < plan_nstages(L, E); ++s)
    if (plan_radix(L, E, s) != E) return false;
  return L / E <= 64 && 64 % (L / E) == 0;
}

template <bool WL> __device__ __forceinline__ void tile_sync() {
  if constexpr (WL) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// tw: LDS table with tw[m*TWS] = exp(-2 pi i m / L) for m in [0, L).
template <typename T, int L, int C, int NT, int DIR, int TWS, int S, bool SRC_LDS, bool RM, bool WL, class LoadF,
          class StoreF>
__device__ __forceinline__ void fft_stage(cx_t<T>* tile, const cx_t<T>* tw, LoadF& load, StoreF& store) {
  using C2 = cx_t<T>;
  constexpr int E = C * L / NT;
  using TI = TileIdx<T, L, C, RM, E>;
  constexpr int RMAX = E < 16 ? E : 16;
  constexpr int NS = plan_nstages(L, RMAX);
  constexpr int R = plan_radix(L, RMAX, S);
  constexpr int P = plan_prefix(L, RMAX, S);
  constexpr int NB = E / R;
  constexpr int STR = L / R;
  constexpr int ROWS = NT / C;
  static_assert(RM || NT % C == 0, "column-major tiles need NT to be a multiple of C");
  static_assert(!WL || (RM && wave_local(L, E)), "wavefront-local exchange needs whole columns per wavefront");
  static_assert(E % R == 0, "per-thread elements must be a multiple of the radix");
  auto bfly = [](int q, int& i, int& col) {
    if constexpr (RM) {
      const int b = threadIdx.x + q * NT;
      i = b % STR;
      col = b / STR;
    } else {
      i = threadIdx.x / C + q * ROWS;
      col = threadIdx.x % C;
    }
  };
  C2 v[E];
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    int i, col;
    bfly(q, i, col);
#pragma unroll
    for (int t = 0; t < R; ++t) {
      const int pos = i + t * STR;
      if constexpr (S == 0) v[q * R + t] = call_load(load, pos, col, q * R + t);
      else v[q * R + t] = tile[TI::at(pos, col)];
    }
  }
  if constexpr (S > 0 || SRC_LDS) tile_sync<WL>();
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    int i, col;
    bfly(q, i, col);
    const int k = i % P;
    if constexpr (S > 0) {
#pragma unroll
      for (int t = 1; t < R; ++t) {
        C2 w = tw[(t * k * (L / (P * R))) * TWS];
        if constexpr (DIR > 0) w.y = -w.y;
        v[q * R + t] = cmul(v[q * R + t], w);
      }
    }
    dftR<R, DIR>(&v[q * R]);
#pragma unroll
    for (int u = 0; u < R; ++u) {
      const int pos = (i - k) * R + k + u * P;
      if constexpr (S == NS - 1) call_store(store, pos, col, v[q * R + u], q * R + u);
      else tile[TI::at(pos, col)] = v[q * R + u];
    }
  }
  if constexpr (S < NS - 1) {
    tile_sync<WL>();
    fft_stage<T, L, C, NT, DIR, TWS, S + 1, SRC_LDS, RM, WL, LoadF, StoreF>(tile, tw, load, store);
  }
}

template <typename T, int L, int C, int NT, int DIR, int TWS, bool SRC_LDS, bool RM = false, bool WL = false,
          class LoadF, class StoreF>
__device__ __forceinline__ void fft_tile(cx_t<T>* tile, const cx_t<T>* tw, LoadF load, StoreF store) {
  tile_sync<WL>();  // the tile may still be read by the previous transform
  fft_stage<T, L, C, NT, DIR, TWS, 0, SRC_LDS, RM, WL, LoadF, StoreF>(tile, tw, load, store);
}

// ---------------------------------------------------------------- split-exchange variant
// Same transform with the LDS exchange between stages done one component
// at a time (real parts, then imaginary parts) through a real-valued tile:
// half the LDS of the complex tile, so twice as many columns fit per CU.
// Column-major layout only; the first stage loads and the last stage stores
// through the caller's functors.
template <typename T, int L, int C, int NT, int DIR, int S, int PL, class StoreF>
__device__ __forceinline__ void fft_stage_split(T* tile, const cx_t<T>* tw, cx_t<T>* v, StoreF& store) {
  using C2 = cx_t<T>;
  constexpr int CP = C + 1;
  constexpr int E = C * L / NT;
  constexpr int RMAX = E < 16 ? E : 16;
  constexpr int NS = plan_nstages(L, RMAX);
  constexpr int R = plan_radix(L, RMAX, S, PL);
  constexpr int P = plan_prefix(L, RMAX, S, PL);
  constexpr int NB = E / R;
  constexpr int ROWS = NT / C;
  const int col = threadIdx.x % C;
  const int i0 = threadIdx.x / C;
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    const int i = i0 + q * ROWS;
    const int k = i % P;
    if constexpr (S > 0) {
#pragma unroll
      for (int t = 1; t < R; ++t) {
        C2 w = tw[t * k * (L / (P * R))];
        if constexpr (DIR > 0) w.y = -w.y;
        v[q * R + t] = cmul(v[q * R + t], w);
      }
    }
    dftR<R, DIR>(&v[q * R]);
  }
  if constexpr (S == NS - 1) {
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const int i = i0 + q * ROWS;
      const int k = i % P;
#pragma unroll
      for (int u = 0; u < R; ++u) call_store(store, (i - k) * R + k + u * P, col, v[q * R + u], q * R + u);
    }
  } else {
    constexpr int R2 = plan_radix(L, RMAX, S + 1, PL);
    constexpr int NB2 = E / R2;
    constexpr int STR2 = L / R2;
    C2 w[E];
#pragma unroll
    for (int part = 0; part < 2; ++part) {
#pragma unroll
      for (int q = 0; q < NB; ++q) {
        const int i = i0 + q * ROWS;
        const int k = i % P;
#pragma unroll
        for (int u = 0; u < R; ++u)
          tile[((i - k) * R + k + u * P) * CP + col] = part ? v[q * R + u].y : v[q * R + u].x;
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < NB2; ++q) {
        const int i = i0 + q * ROWS;
#pragma unroll
        for (int t = 0; t < R2; ++t) {
          const T x = tile[(i + t * STR2) * CP + col];
          if (part) w[q * R2 + t].y = x;
          else w[q * R2 + t].x = x;
        }
      }
      __syncthreads();
    }
    fft_stage_split<T, L, C, NT, DIR, S + 1, PL, StoreF>(tile, tw, w, store);
  }
}

// The first-stage inputs of fft_tile_split, loaded into registers through the
// caller's functor; split_transform finishes the transform from them.  Apart,
// they let a caller hold the next tile's inputs while it transforms this one.
template <typename T, int L, int C, int NT, int PL = 0, class LoadF>
__device__ __forceinline__ void split_load(cx_t<T>* v, LoadF load) {
  constexpr int E = C * L / NT;
  constexpr int RMAX = E < 16 ? E : 16;
  constexpr int R = plan_radix(L, RMAX, 0, PL);
  constexpr int NB = E / R;
  constexpr int STR = L / R;
  constexpr int ROWS = NT / C;
  const int col = threadIdx.x % C;
  const int i0 = threadIdx.x / C;
#pragma unroll
  for (int q = 0; q < NB; ++q)
#pragma unroll
    for (int t = 0; t < R; ++t) v[q * R + t] = call_load(load, i0 + q * ROWS + t * STR, col, q * R + t);
}

template <typename T, int L, int C, int NT, int DIR, int PL = 0, class StoreF>
__device__ __forceinline__ void split_transform(T* tile, const cx_t<T>* tw, cx_t<T>* v, StoreF store) {
  __syncthreads();  // the tile may still be read by the previous transform
  fft_stage_split<T, L, C, NT, DIR, 0, PL, StoreF>(tile, tw, v, store);
}

template <typename T, int L, int C, int NT, int DIR, int PL = 0, class LoadF, class StoreF>
__device__ __forceinline__ void fft_tile_split(T* tile, const cx_t<T>* tw, LoadF load, StoreF store) {
  cx_t<T> v[C * L / NT];
  split_load<T, L, C, NT, PL>(v, load);
  split_transform<T, L, C, NT, DIR, PL>(tile, tw, v, store);
}

// Element slots of the first stage of fft_tile_split: the (pos, col) that
// slot e of this thread is loaded from.
template <typename T, int L, int C, int NT, int PL = 0>
__device__ __forceinline__ void split_slot(int e, int& pos, int& col) {
  constexpr int E = C * L / NT;
  constexpr int RMAX = E < 16 ? E : 16;
  constexpr int R = plan_radix(L, RMAX, 0, PL);
  constexpr int STR = L / R;
  constexpr int ROWS = NT / C;
  col = threadIdx.x % C;
  pos = threadIdx.x / C + (e / R) * ROWS + (e % R) * STR;
}

// Input slots of the first stage of fft_tile: the (pos, col) that slot e of
// this thread is loaded from.
template <typename T, int L, int C, int NT, bool RM>
__device__ __forceinline__ void first_slot(int e, int& pos, int& col) {
  constexpr int E = C * L / NT;
  constexpr int RMAX = E < 16 ? E : 16;
  constexpr int R = plan_radix(L, RMAX, 0);
  constexpr int STR = L / R;
  constexpr int ROWS = NT / C;
  const int q = e / R, t = e % R;
  int i;
  if constexpr (RM) {
    const int b = threadIdx.x + q * NT;
    i = b % STR;
    col = b / STR;
  } else {
    i = threadIdx.x / C + q * ROWS;
    col = threadIdx.x % C;
  }
  pos = i + t * STR;
}

// Output slots of the last stage of fft_tile: the (pos, col) that slot e of
// this thread is stored to.
template <typename T, int L, int C, int NT, bool RM>
__device__ __forceinline__ void last_slot(int e, int& pos, int& col) {
  constexpr int E = C * L / NT;
  constexpr int RMAX = E < 16 ? E : 16;
  constexpr int NS = plan_nstages(L, RMAX);
  constexpr int R = plan_radix(L, RMAX, NS - 1);
  constexpr int P = plan_prefix(L, RMAX, NS - 1);
  constexpr int STR = L / R;
  constexpr int ROWS = NT / C;
  const int q = e / R, u = e % R;
  int i;
  if constexpr (RM) {
    const int b = threadIdx.x + q * NT;
    i = b % STR;
    col = b / STR;
  } else {
    i = threadIdx.x / C + q * ROWS;
    col = threadIdx.x % C;
  }
  const int k = i % P;
  pos = (i - k) * R + k + u * P;
}

// Copy a global twiddle table (exp(-2 pi i m / n), m < n) into LDS.
template <typename C2>
__device__ __forceinline__ void load_twiddles(C2* dst, const C2* src, int n, int nt) {
  for (int m = threadIdx.x; m < n; m += nt) dst[m] = src[m];
}

// ---------------------------------------------------------------- mixed-radix transforms
// Lengths that are not powers of two (mesh sizes 48, 96, 160, 384, 768, ...) run through a
// Stockham transform with a run-time plan of radices 8, 4, 2, 3, 5 and 7.  C independent
// vectors are interleaved in LDS, element (pos, c) at pos * (C + 1) + c; each stage reads one
// buffer and writes the other (the caller's two buffers), so a thread holds one butterfly at a
// time and the plan needs no per-thread element count divisible by every radix.  Stage s of
// radix R with P = the product of the earlier radices computes, for every butterfly
// i in [0, L/R) with k = i mod P:  v_t = x[i + t L/R] w^(t k L/(P R)),  y = DFT_R(v),
// x'[(i - k) R + k + u P] = y_u  (the same recurrence as the power-of-two tiles above).
struct GPlan {
  int L, ns;
  int r[12];
};

// radices for a length L whose prime factors are 2, 3, 5 and 7: eights first, then a four
// or a two, then threes, fives and sevens (host)
inline GPlan make_gplan(int L) {
  GPlan p{};
  p.L = L;
  int n = L;
  while (n % 8 == 0 && p.ns < 12) p.r[p.ns++] = 8, n /= 8;
  if (n % 4 == 0) p.r[p.ns++] = 4, n /= 4;
  if (n % 2 == 0) p.r[p.ns++] = 2, n /= 2;
  const int odd[3] = {3, 5, 7};
  for (int q : odd)
    while (n % q == 0 && p.ns < 12) p.r[p.ns++] = q, n /= q;
  if (n != 1) p.ns = 0;  // not 7-smooth: no plan
  return p;
}

// cos and sin of 2 pi m / R for R = 3, 5, 7
template <int R> __device__ __forceinline__ double cos_r(int m) {
  if constexpr (R == 3) return -0.5;
  else if constexpr (R == 5) {
    return (m == 1 || m == 4) ? 0.309016994374947424102293417182819059 : -0.809016994374947424102293417182819059;
  } else {
    return (m == 1 || m == 6) ? 0.623489801858733530525004884004239810
           : (m == 2 || m == 5) ? -0.222520933956314404288902564496794759
                                : -0.900968867902419126236102319507445051;
  }
}
template <int R> __device__ __forceinline__ double sin_r(int m) {
  const double s = m < R - m ? 1.0 : -1.0;
  const int a = m < R - m ? m : R - m;
  if constexpr (R == 3) return s * 0.866025403784438646763723170752936183;
  else if constexpr (R == 5) {
    return s * (a == 1 ? 0.951056516295153572116439333379382143 : 0.587785252292473129168705954639072769);
  } else {
    return s * (a == 1 ? 0.781831482468029808708444526674057750
                : a == 2 ? 0.974927912181823607018131682993931217
                         : 0.433883739117558120475768332848358754);
  }
}

// DFT of odd prime length R by its definition, y_u = sum_t v_t exp(DIR 2 pi i u t / R)
template <int R, int DIR, typename C> __device__ __forceinline__ void dftp(C* v) {
  using T = decltype(C::x);
  C y[R];
#pragma unroll
  for (int u = 0; u < R; ++u) {
    C acc = v[0];
#pragma unroll
    for (int t = 1; t < R; ++t) {
      const int m = (u * t) % R;
      if (m == 0) {
        acc = cadd(acc, v[t]);
      } else {
        const T c = (T)cos_r<R>(m), s = (T)(DIR * sin_r<R>(m));
        acc.x = fma(v[t].x, c, fma(-v[t].y, s, acc.x));
        acc.y = fma(v[t].x, s, fma(v[t].y, c, acc.y));
      }
    }
    y[u] = acc;
  }
#pragma unroll
  for (int u = 0; u < R; ++u) v[u] = y[u];
}

template <int R, int DIR, typename C> __device__ __forceinline__ void dft_any(C* v) {
  if constexpr (R == 3 || R == 5 || R == 7) dftp<R, DIR>(v);
  else dftR<R, DIR>(v);
}

// One stage of radix R over C interleaved vectors: src -> dst.  tw[m * tws] = exp(-2 pi i m / L).
template <typename T, int R, int DIR, int C>
__device__ __forceinline__ void gstage(const cx_t<T>* src, cx_t<T>* dst, const cx_t<T>* tw, int tws, int L, int P,
                                       int nt) {
  using C2 = cx_t<T>;
  constexpr int CP = C + 1;
  const int S = L / R, tm = L / (P * R);
  // k = i mod P without an integer division: floor((i + 1/2) / P) in FP32 is exact for
  // i, P < 2^11 (the quotient stays 1/(2P) away from every integer)
  const float rp = 1.0f / (float)P;
  for (int q = threadIdx.x; q < C * S; q += nt) {
    const int c = q % C, i = q / C;
    const int k = i - P * (int)(((float)i + 0.5f) * rp);
    C2 v[R];
#pragma unroll
    for (int t = 0; t < R; ++t) v[t] = src[(i + t * S) * CP + c];
    if (P > 1) {
#pragma unroll
      for (int t = 1; t < R; ++t) {
        C2 w = tw[t * k * tm * tws];
        if constexpr (DIR > 0) w.y = -w.y;
        v[t] = cmul(v[t], w);
      }
    }
    dft_any<R, DIR>(v);
#pragma unroll
    for (int u = 0; u < R; ++u) dst[((i - k) * R + k + u * P) * CP + c] = v[u];
  }
}

// The whole transform, starting in a (already written and synchronised); returns the buffer
// that holds the result (a or b), synchronised.
template <typename T, int DIR, int C>
__device__ cx_t<T>* gfft(cx_t<T>* a, cx_t<T>* b, const cx_t<T>* tw, int tws, const GPlan& pl, int nt) {
  int P = 1;
  for (int s = 0; s < pl.ns; ++s) {
    const int R = pl.r[s];
    switch (R) {
      case 2: gstage<T, 2, DIR, C>(a, b, tw, tws, pl.L, P, nt); break;
      case 3: gstage<T, 3, DIR, C>(a, b, tw, tws, pl.L, P, nt); break;
      case 4: gstage<T, 4, DIR, C>(a, b, tw, tws, pl.L, P, nt); break;
      case 5: gstage<T, 5, DIR, C>(a, b, tw, tws, pl.L, P, nt); break;
      case 7: gstage<T, 7, DIR, C>(a, b, tw, tws, pl.L, P, nt); break;
      default: gstage<T, 8, DIR, C>(a, b, tw, tws, pl.L, P, nt); break;
    }
    __syncthreads();
    P *= R;
    cx_t<T>* t = a;
    a = b;
    b = t;
  }
  return a;
}
