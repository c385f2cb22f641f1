// fft_device.hpp -- register/LDS Stockham line FFTs for gfx950 (CDNA4).
//
// A "line" is one 1-D transform of length N.  Each thread of a line owns E
// complex values; with TL = N/E threads per line, thread t holds line
// positions t + TL*e (e = 0..E-1) both on entry and on exit, in natural
// order.  That strided ownership is what makes the first and last pass talk
// to HBM directly with coalesced accesses: for a column tile the lanes of a
// wave are (column, t) pairs, for a row the lanes are consecutive t.
//
// Passes are Stockham radix-R steps (R in {2,3,4,5,6,8,12,16}); between passes
// the values are exchanged through LDS (re and im as separate planes, so an
// exchange can also run one plane at a time in half the LDS).  The last pass
// writes its butterfly outputs straight back into the owning registers, so a
// transform with P passes needs P-1 exchanges.
//
// Twiddles for pass p are read from a device table laid out as
// tw[off_p + k*(R-1) + (q-1)] = exp(-2*pi*i*q*k/(Ns*R)), k < Ns, 1 <= q < R,
// written by the host in long double and rounded once.  The inverse uses
// the conjugates.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace pfc_amd {

template <typename T> struct cplx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ cplx<T> c_add(cplx<T> a, cplx<T> b) {
  return {a.x + b.x, a.y + b.y};
}
template <typename T>
__device__ __forceinline__ cplx<T> c_sub(cplx<T> a, cplx<T> b) {
  return {a.x - b.x, a.y - b.y};
}
template <typename T>
__device__ __forceinline__ cplx<T> c_mul(cplx<T> a, cplx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * conj(b)
template <typename T>
__device__ __forceinline__ cplx<T> c_mulc(cplx<T> a, cplx<T> b) {
  return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <typename T>
__device__ __forceinline__ cplx<T> c_scale(cplx<T> a, T s) {
  return {a.x * s, a.y * s};
}
template <typename T> __device__ __forceinline__ cplx<T> c_conj(cplx<T> a) {
  return {a.x, -a.y};
}
// multiply by -i (DIR<0) or +i (DIR>0)
template <int DIR, typename T>
__device__ __forceinline__ cplx<T> c_rot(cplx<T> a) {
  if constexpr (DIR < 0)
    return {a.y, -a.x};
  else
    return {-a.y, a.x};
}

// ---------------------------------------------------------------------------
// Constant twiddles exp(DIR * 2*pi*i*m/R) for the small in-register DFTs.
// cos/sin of multiples of pi/8, pi/6 and 2*pi/5 to full double precision.
// ---------------------------------------------------------------------------
struct cs {
  double c, s;
};

constexpr cs unit_root(int m, int R) {
  // returns (cos(2 pi m/R), sin(2 pi m/R)) for R in {2,3,4,5,6,8,12,16}
  m = ((m % R) + R) % R;
  constexpr double C8 = 0.92387953251128675613; // cos(pi/8)
  constexpr double S8 = 0.38268343236508977173; // sin(pi/8)
  constexpr double H2 = 0.70710678118654752440; // sqrt(2)/2
  constexpr double H3 = 0.86602540378443864676; // sqrt(3)/2
  constexpr double C5a = 0.30901699437494742410, S5a = 0.95105651629515357212;
  constexpr double C5b = -0.80901699437494742410, S5b = 0.58778525229247312917;
  if (R == 16) {
    constexpr cs t[16] = {{1, 0},     {C8, S8},   {H2, H2},   {S8, C8},
                          {0, 1},     {-S8, C8},  {-H2, H2},  {-C8, S8},
                          {-1, 0},    {-C8, -S8}, {-H2, -H2}, {-S8, -C8},
                          {0, -1},    {S8, -C8},  {H2, -H2},  {C8, -S8}};
    return t[m];
  }
  if (R == 8) return unit_root(2 * m, 16);
  if (R == 4) return unit_root(4 * m, 16);
  if (R == 2) return unit_root(8 * m, 16);
  if (R == 12) {
    constexpr cs t[12] = {{1, 0},     {H3, 0.5},   {0.5, H3},   {0, 1},
                          {-0.5, H3}, {-H3, 0.5},  {-1, 0},     {-H3, -0.5},
                          {-0.5, -H3}, {0, -1},    {0.5, -H3},  {H3, -0.5}};
    return t[m];
  }
  if (R == 6) return unit_root(2 * m, 12);
  if (R == 3) return unit_root(4 * m, 12);
  if (R == 5) {
    constexpr cs t[5] = {{1, 0}, {C5a, S5a}, {C5b, S5b}, {C5b, -S5b},
                         {C5a, -S5a}};
    return t[m];
  }
  return {1, 0};
}

// a * exp(DIR*2*pi*i*m/R), exact shortcuts on the axes.
template <int DIR, int R, int M, typename T>
__device__ __forceinline__ cplx<T> c_twc(cplx<T> a) {
  constexpr int m = ((M % R) + R) % R;
  if constexpr (m == 0) {
    return a;
  } else if constexpr (4 * m == R) { // exp(DIR*i*pi/2)
    return c_rot<DIR>(a);
  } else if constexpr (2 * m == R) {
    return {-a.x, -a.y};
  } else if constexpr (4 * m == 3 * R) {
    return c_rot<-DIR>(a);
  } else {
    constexpr cs w = unit_root(m, R);
    const T c = (T)w.c;
    const T s = (T)(DIR * w.s);
    return {a.x * c - a.y * s, a.x * s + a.y * c};
  }
}

// ---------------------------------------------------------------------------
// In-register DFTs, natural order in and out, unnormalised.
// DIR = -1 forward (exp(-2 pi i nk/R)), +1 inverse.
// ---------------------------------------------------------------------------
template <int R, int DIR, typename T> struct Dft;

template <int DIR, typename T> struct Dft<1, DIR, T> {
  __device__ __forceinline__ static void run(cplx<T> *) {}
};

template <int DIR, typename T> struct Dft<2, DIR, T> {
  __device__ __forceinline__ static void run(cplx<T> *a) {
    cplx<T> t = a[0];
    a[0] = c_add(t, a[1]);
    a[1] = c_sub(t, a[1]);
  }
};

template <int DIR, typename T> struct Dft<4, DIR, T> {
  __device__ __forceinline__ static void run(cplx<T> *a) {
    cplx<T> t0 = c_add(a[0], a[2]), t1 = c_sub(a[0], a[2]);
    cplx<T> t2 = c_add(a[1], a[3]), t3 = c_rot<DIR>(c_sub(a[1], a[3]));
    a[0] = c_add(t0, t2);
    a[2] = c_sub(t0, t2);
    a[1] = c_add(t1, t3);
    a[3] = c_sub(t1, t3);
  }
};

template <int DIR, typename T> struct Dft<3, DIR, T> {
  __device__ __forceinline__ static void run(cplx<T> *a) {
    const T h3 = (T)0.86602540378443864676;
    cplx<T> s = c_add(a[1], a[2]);
    cplx<T> d = c_sub(a[1], a[2]);
    cplx<T> m = {a[0].x - (T)0.5 * s.x, a[0].y - (T)0.5 * s.y};
    // DIR*i*h3*d
    cplx<T> r = c_rot<DIR>(cplx<T>{h3 * d.x, h3 * d.y});
    a[0] = c_add(a[0], s);
    a[1] = c_add(m, r);
    a[2] = c_sub(m, r);
  }
};

// Generic O(R^2) DFT with constant roots (used for radix 5).
template <int DIR, typename T> struct Dft<5, DIR, T> {
  template <int K, int Q>
  __device__ __forceinline__ static void acc(cplx<T> &o, const cplx<T> *in) {
    cplx<T> v = c_twc<DIR, 5, K * Q>(in[Q]);
    o = c_add(o, v);
    if constexpr (Q + 1 < 5) acc<K, Q + 1>(o, in);
  }
  template <int K>
  __device__ __forceinline__ static void out(cplx<T> *o, const cplx<T> *in) {
    o[K] = in[0];
    acc<K, 1>(o[K], in);
    if constexpr (K + 1 < 5) out<K + 1>(o, in);
  }
  __device__ __forceinline__ static void run(cplx<T> *a) {
    cplx<T> in[5] = {a[0], a[1], a[2], a[3], a[4]};
    out<0>(a, in);
  }
};

// Two-level P x Q decomposition (R = P*Q): n = n1 + P*n2, k = Q*k1 + k2.
template <int P, int Q, int DIR, typename T> struct DftPQ {
  static constexpr int R = P * Q;
  template <int N1, int K2>
  __device__ __forceinline__ static void tw(cplx<T> *y) {
    y[N1 * Q + K2] = c_twc<DIR, R, N1 * K2>(y[N1 * Q + K2]);
    if constexpr (K2 + 1 < Q) tw<N1, K2 + 1>(y);
    else if constexpr (N1 + 1 < P) tw<N1 + 1, 0>(y);
  }
  __device__ __forceinline__ static void run(cplx<T> *a) {
    cplx<T> y[R];
#pragma unroll
    for (int n1 = 0; n1 < P; ++n1) {
      cplx<T> t[Q];
#pragma unroll
      for (int n2 = 0; n2 < Q; ++n2) t[n2] = a[n1 + P * n2];
      Dft<Q, DIR, T>::run(t);
#pragma unroll
      for (int k2 = 0; k2 < Q; ++k2) y[n1 * Q + k2] = t[k2];
    }
    tw<0, 0>(y);
#pragma unroll
    for (int k2 = 0; k2 < Q; ++k2) {
      cplx<T> t[P];
#pragma unroll
      for (int n1 = 0; n1 < P; ++n1) t[n1] = y[n1 * Q + k2];
      Dft<P, DIR, T>::run(t);
#pragma unroll
      for (int k1 = 0; k1 < P; ++k1) a[Q * k1 + k2] = t[k1];
    }
  }
};

template <int DIR, typename T> struct Dft<8, DIR, T> {
  __device__ __forceinline__ static void run(cplx<T> *a) {
    DftPQ<2, 4, DIR, T>::run(a);
  }
};
template <int DIR, typename T> struct Dft<16, DIR, T> {
  __device__ __forceinline__ static void run(cplx<T> *a) {
    DftPQ<4, 4, DIR, T>::run(a);
  }
};
template <int DIR, typename T> struct Dft<6, DIR, T> {
  __device__ __forceinline__ static void run(cplx<T> *a) {
    DftPQ<2, 3, DIR, T>::run(a);
  }
};
template <int DIR, typename T> struct Dft<12, DIR, T> {
  __device__ __forceinline__ static void run(cplx<T> *a) {
    DftPQ<4, 3, DIR, T>::run(a);
  }
};

// ---------------------------------------------------------------------------
// Compile-time transform plans.
// ---------------------------------------------------------------------------
template <int... Rs> struct RadixList {};

// XP_: the butterflies (twiddle products and the in-register DFTs) run in
// double whatever the value type, from a twiddle table of double entries;
// values are rounded back to the value type after each radix pass.  Used by
// single-precision fields whose transforms must keep FP64 arithmetic error
// (the FP64-butterfly FP32-mixed step).
template <int N_, int E_, typename List, bool XP_ = false> struct FftPlan;
template <int N_, int E_, int... Rs, bool XP_> struct FftPlan<N_, E_, RadixList<Rs...>, XP_> {
  static constexpr int N = N_;
  static constexpr int E = E_;
  static constexpr int TL = N_ / E_; // threads per line
  static constexpr int NPASS = sizeof...(Rs);
  static constexpr bool XP = XP_;
  // every radix a power of two (the XOR-swizzled row exchanges need it)
  static constexpr bool POW2 = (((Rs & (Rs - 1)) == 0) && ...);
  using List = RadixList<Rs...>;
  static_assert((Rs * ... * 1) == N_, "radices must multiply to N");
  static_assert(((E_ % Rs == 0) && ...), "every radix must divide E");
  // twiddle table length (complex entries)
  static constexpr int twiddle_len() {
    int ns = 1, len = 0;
    ((len += (ns > 1 ? ns * (Rs - 1) : 0), ns *= Rs), ...);
    return len;
  }
};

// LDS view for one line: element pos of this line lives at
// re[pos*spos + off], im[...] (im plane follows re plane at `plane`).
template <typename T> struct LdsLine {
  T *base;  // start of the re plane for the whole workgroup
  int off;  // this line's offset within a plane
  int spos; // stride between consecutive positions
  int plane; // distance (elements) from the re plane to the im plane (0: split)
  __device__ __forceinline__ int idx(int pos) const {
    return pos * spos + off;
  }
};

// LDS padding: one pad position every 16 positions breaks the power-of-two
// strides of the Stockham write pattern across LDS banks (used by rows and
// by column tiles, whose L columns are interleaved per position).
__host__ __device__ constexpr int row_padded(int n) { return n + n / 16; }

template <bool ROWPAD>
__device__ __forceinline__ int pad_pos(int pos) {
  if constexpr (ROWPAD)
    return pos + (pos >> 4);
  else
    return pos;
}

// Ordering of the LDS exchanges of one line.  WAVE: the line's threads lie
// in one wavefront (row layouts with TL <= 64 dividing 64), whose LDS
// instructions execute in program order, so a wavefront barrier (which also
// keeps the compiler from moving LDS accesses across it) is enough and the
// other lines of the block need not wait.  Otherwise a block barrier.
template <bool WAVE> __device__ __forceinline__ void lds_sync() {
  if constexpr (WAVE) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// Row layouts (a line's TL threads consecutive in the block) whose lines fit
// in one wavefront.
template <typename Plan> constexpr bool row_wave() {
  return Plan::TL <= 64 && 64 % Plan::TL == 0;
}

// One Stockham pass (radix R, current span NS) followed, when more passes
// remain, by an LDS exchange.
template <typename T, typename Plan, int DIR, bool SPLIT, bool ROWPAD, bool WAVE, int NS,
          int TWOFF, int R, int... Rest>
struct Pass {
  static constexpr int N = Plan::N, E = Plan::E, TL = Plan::TL;
  static constexpr int NB = E / R; // butterflies per thread
  static constexpr bool LAST = (NS * R == N);

  // arithmetic type of the butterflies (double for XP plans)
  using C = std::conditional_t<Plan::XP, double, T>;

  __device__ __forceinline__ static void run(cplx<T> (&v)[E], int t,
                                             const LdsLine<T> &lds,
                                             const cplx<T> *__restrict__ tw) {
    // butterflies; element q of butterfly b is v[b + NB*q]
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      cplx<C> a[R];
#pragma unroll
      for (int q = 0; q < R; ++q) a[q] = cplx<C>{(C)v[b + NB * q].x, (C)v[b + NB * q].y};
      const int j = t + TL * b;
      const int k = j % NS;
      if constexpr (NS > 1) {
        // XP plans read a table of double entries through the same pointer
        const cplx<C> *w = reinterpret_cast<const cplx<C> *>(tw) + TWOFF + k * (R - 1);
        if constexpr (std::is_same_v<T, double> && R == 16) {
          // FP64 fields, radix 16: only w^(2^i) come from the table (4 loads
          // instead of 15); w^q and w^(q + 8), q < 8, are products of those,
          // formed as they are used, so that few twiddles are live at once
          // (each within a few roundings of the table's entry, far below the
          // step's tolerance)
          constexpr int H = R / 2;
          cplx<C> p[4]; // p[i] = w^(2^i)
#pragma unroll
          for (int i = 0; (1 << i) < R; ++i) p[i] = w[(1 << i) - 1];
          const cplx<C> wh = p[H == 2 ? 1 : (H == 4 ? 2 : 3)];
#pragma unroll
          for (int q = 1; q < H; ++q) {
            cplx<C> wq = p[0];
            bool first = true;
#pragma unroll
            for (int i = 0; (1 << i) < H; ++i)
              if (q & (1 << i)) {
                wq = first ? p[i] : c_mul(wq, p[i]);
                first = false;
              }
            const cplx<C> wq2 = c_mul(wq, wh);
            a[q] = (DIR < 0) ? c_mul(a[q], wq) : c_mulc(a[q], wq);
            a[q + H] = (DIR < 0) ? c_mul(a[q + H], wq2) : c_mulc(a[q + H], wq2);
          }
          a[H] = (DIR < 0) ? c_mul(a[H], wh) : c_mulc(a[H], wh);
        } else {
#pragma unroll
          for (int q = 1; q < R; ++q) {
            cplx<C> wq = w[q - 1];
            a[q] = (DIR < 0) ? c_mul(a[q], wq) : c_mulc(a[q], wq);
          }
        }
      }
      Dft<R, DIR, C>::run(a);
      // outputs go back to slot b + NB*q (parked there until the exchange
      // unless this is the last pass)
#pragma unroll
      for (int q = 0; q < R; ++q) v[b + NB * q] = cplx<T>{(T)a[q].x, (T)a[q].y};
    }
    if constexpr (!LAST) {
      exchange(v, t, lds);
      Pass<T, Plan, DIR, SPLIT, ROWPAD, WAVE, NS * R,
           TWOFF + (NS > 1 ? NS * (R - 1) : 0), Rest...>::run(v, t, lds, tw);
    }
  }

  // Destination position of butterfly b, output q.
  __device__ __forceinline__ static int dst_pos(int t, int b, int q) {
    const int j = t + TL * b;
    return (j / NS) * NS * R + (j % NS) + q * NS;
  }

  // LDS slot of line position pos in this exchange.  The LDS serves a
  // wavefront's 4-byte accesses in halves of 32 lanes over 32 banks, its
  // 8-byte accesses in quarters of 16 lanes over 16 bank pairs.  One pad
  // slot per 16 positions suits the 8-byte planes: a quarter reads 16
  // consecutive positions and the writes of stride R <= 16 spread (only the
  // middle exchanges of R <= 8 plans pair up).  For 4-byte planes of
  // power-of-two plans whose line fills at least half a wavefront, where a
  // half reads 32 consecutive positions (aligned to 32) and writes 32 / NS
  // groups of NS consecutive positions A = NS*R apart:
  //  - NS == 1 or NS >= 32: one pad slot per 32 positions (writes of stride R
  //    <= 16 and contiguous writes spread; reads shift as a block);
  //  - R == 16 (A >= 32): NS pad slots per A positions, which moves the
  //    groups NS banks apart.
  // Both stay within row_padded(N) and give every read and write address as
  // a per-lane part plus a compile-time offset.  The remaining exchanges
  // (1 < NS < 32 with R <= 8) keep one pad slot per 16 positions: an XOR
  // swizzle would spread them as well, but its write addresses differ per
  // output in their low bits, and the compiler then holds them in registers
  // across the transforms of a pass.  Column layouts are unpadded.
  static constexpr bool SWZ = ROWPAD && sizeof(T) == 4 && Plan::POW2 && TL >= 32 && N % 32 == 0;
  static constexpr int SWZ_KIND = (NS == 1 || NS >= 32) ? 0 : (R == 16 ? 1 : 2);
  __device__ __forceinline__ static int slot(int pos) {
    if constexpr (!SWZ || SWZ_KIND == 2)
      return pad_pos<ROWPAD>(pos);
    else if constexpr (SWZ_KIND == 0)
      return pos + (pos >> 5);
    else
      return pos + (pos / (NS * R)) * NS;
  }
  // slot(t + TL*e), t < TL, as a per-lane part plus a per-e constant (TL is
  // a multiple of 32 where SWZ holds), so that the reads of all e share one
  // address register and take the rest as an immediate offset.
  __device__ __forceinline__ static int read_slot(int t, int e) {
    if constexpr (ROWPAD && TL % 16 == 0 && (!SWZ || SWZ_KIND == 2)) {
      return t + (t >> 4) + e * (TL + TL / 16); // pad_pos(t + TL*e), split
    } else if constexpr (!SWZ || SWZ_KIND == 2) {
      return pad_pos<ROWPAD>(t + TL * e);
    } else if constexpr (SWZ_KIND == 0) {
      return t + (t >> 5) + e * (TL + TL / 32);
    } else {
      constexpr int A = NS * R;
      if constexpr (TL % A == 0)
        return t + (t / A) * NS + e * (TL + (TL / A) * NS);
      else // A a multiple of TL: t + TL*e stays within one block of A
        return t + TL * e + ((TL * e) / A) * NS;
    }
  }

  // slot(dst_pos(t, b, q)) as a per-(lane, b) part plus q times a constant.
  __device__ __forceinline__ static int write_slot(int t, int b, int q) {
    if constexpr (ROWPAD && (!SWZ || SWZ_KIND == 2) &&
                  ((NS == 1 && R <= 16 && (R & (R - 1)) == 0) || (NS > 1 && NS % 16 == 0))) {
      // pad_pos(d0 + q*NS), split: d0 = R*j with R dividing 16 (NS == 1; so
      // d0 mod 16 + q < 16), or q*NS a multiple of 16
      const int d0 = dst_pos(t, b, 0);
      return d0 + (d0 >> 4) + q * (NS == 1 ? 1 : NS + NS / 16);
    } else if constexpr (!SWZ || SWZ_KIND == 2) {
      return pad_pos<ROWPAD>(dst_pos(t, b, q));
    } else {
      const int d0 = dst_pos(t, b, 0);
      if constexpr (SWZ_KIND == 1) // dst / A = d0 / A for every q
        return d0 + (d0 / (NS * R)) * NS + q * NS;
      else if constexpr (NS == 1) // d0 = R*j: d0 mod 32 + q < 32
        return d0 + (d0 >> 5) + q;
      else // NS >= 32: q*NS is a multiple of 32
        return d0 + (d0 >> 5) + q * (NS + NS / 32);
    }
  }

  __device__ __forceinline__ static void exchange(cplx<T> (&v)[E], int t,
                                                  const LdsLine<T> &lds) {
    if constexpr (SPLIT) {
      // re plane, then im plane, through the same LDS region
#pragma unroll
      for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int q = 0; q < R; ++q)
          lds.base[lds.idx(write_slot(t, b, q))] = v[b + NB * q].x;
      lds_sync<WAVE>();
#pragma unroll
      for (int e = 0; e < E; ++e)
        v[e].x = lds.base[lds.idx(read_slot(t, e))];
      lds_sync<WAVE>();
#pragma unroll
      for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int q = 0; q < R; ++q)
          lds.base[lds.idx(write_slot(t, b, q))] = v[b + NB * q].y;
      lds_sync<WAVE>();
#pragma unroll
      for (int e = 0; e < E; ++e)
        v[e].y = lds.base[lds.idx(read_slot(t, e))];
      lds_sync<WAVE>();
    } else {
#pragma unroll
      for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int q = 0; q < R; ++q) {
          const int i = lds.idx(write_slot(t, b, q));
          lds.base[i] = v[b + NB * q].x;
          lds.base[i + lds.plane] = v[b + NB * q].y;
        }
      lds_sync<WAVE>();
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int i = lds.idx(read_slot(t, e));
        v[e].x = lds.base[i];
        v[e].y = lds.base[i + lds.plane];
      }
      lds_sync<WAVE>();
    }
  }
};

// One line transform.  WAVE (see lds_sync): the exchanges wait for the
// line's own wavefront only; the caller then separates any use of the LDS
// by other lines, before or after, with a block barrier.
template <typename T, typename Plan, int DIR, bool SPLIT, bool ROWPAD, bool WAVE = false,
          typename List = typename Plan::List>
struct LineFft;
template <typename T, typename Plan, int DIR, bool SPLIT, bool ROWPAD, bool WAVE, int... Rs>
struct LineFft<T, Plan, DIR, SPLIT, ROWPAD, WAVE, RadixList<Rs...>> {
  __device__ __forceinline__ static void run(cplx<T> (&v)[Plan::E], int t,
                                             const LdsLine<T> &lds,
                                             const cplx<T> *__restrict__ tw) {
    Pass<T, Plan, DIR, SPLIT, ROWPAD, WAVE, 1, 0, Rs...>::run(v, t, lds, tw);
  }
};

} // namespace pfc_amd
