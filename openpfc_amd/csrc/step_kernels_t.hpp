// step_kernels_t.hpp -- the fused PFC step on transposed spectral layouts.
//
// Every pass reads or writes whole contiguous lines, one line per TL threads
// (one wave for the 512- and 1024-point plans), and the change of layout
// between passes goes through an LDS tile once per block:
//
//   psi  [z][y][x]   real, x contiguous           (caller's field)
//   T1   [z][kx][y]  complex, y contiguous        (after forward x; inverse y)
//   T2   [kx][ky][z] complex, z contiguous        (after forward y; z pass)
//
//   k_tx_fwd  psi -> T1 (Psi and N):  rows of psi, transposed store
//             (k_tx_fwd2: the same with two nx/2-point R2Cs)
//   k_ty_fwd  T1 -> T2 (both fields): y lines of T1, transposed store
//   k_tz      T2 -> T2 (in place):    z lines; forward both, ETD combine, inverse
//   k_ty_inv  T2 -> T1:               transposed load, y lines of T1
//   k_tx_inv  T1 -> psi:              transposed load, C2R rows of psi
//
// The Tungsten mean-field step runs nine passes on the same layouts (A for
// Psi, B for the mean field and then N):
//   k_tx_r2c psi -> A1 ; k_ty_fwd A1 -> A2 ; k_tz(ZM_TMF) A2 -> A2, B2 ;
//   k_ty_inv B2 -> B1 ; k_tx_tung B1 -> B1 (inverse x, N(psi, psi_MF),
//   forward x) ; k_ty_fwd B1 -> B2 ; k_tz(ZM_TCOMB) A2, B2 -> A2 ;
//   k_ty_inv A2 -> A1 ; k_tx_inv A1 -> psi
// for 3R + 18C bytes (R a real field, C a half spectrum).
//
// A block handles G lines that are adjacent in the transposed dimension, so
// the strided side of each pass moves G*sizeof(complex) contiguous bytes
// (128 bytes for G = 8 in double).  Algorithmic traffic is the same as the
// natural-layout step: R + 2C, 4C, 3C, 2C, C + R.
#pragma once

#include "fft_kernels.hpp"

namespace pfc_amd {

// LDS plane size (elements) of G lines of an N-point plan, row layout.
template <typename Plan, int G> constexpr int t_plane() {
  return G * row_padded(Plan::N);
}

// Transpose tile: positions x G lines, tile_pitch<G, T>() slots per position.
// A line is written position-major (consecutive positions of one line
// across the lanes) and the tile is read row-major (S / G positions x G
// lines per group of S lanes).  The LDS serves a wavefront's 4-byte accesses
// in halves of 32 lanes over 32 banks and its 8-byte accesses in quarters of
// 16 lanes over 16 bank pairs, so S = 128 / sizeof(T) lanes must fall on S
// distinct slots.  For power-of-two G <= S the line index is XORed with
// (pos / (S / G)) mod G within a pitch of G, which does that for both
// patterns (a padding column, pitch G + 1, spreads the writes, but the reads
// of S / G rows then wrap onto the first); other G keep the padding column.
template <int G, typename T> constexpr bool tile_swizzled() {
  return G <= 128 / (int)sizeof(T) && (G & (G - 1)) == 0;
}
template <int G, typename T> constexpr int tile_pitch() {
  return tile_swizzled<G, T>() ? G : G + 1;
}
template <int G, typename T> __device__ __forceinline__ int tile_idx(int pos, int g) {
  constexpr int S = 128 / (int)sizeof(T);
  if constexpr (tile_swizzled<G, T>())
    return pos * G + (g ^ ((pos / (S / G)) & (G - 1)));
  else
    return pos * (G + 1) + g;
}

// Fetch, for every owned element k, the value at the mirror position (L - k,
// 0 for k = 0) of the same line through LDS; one plane at a time when SPLIT
// (half the LDS).
template <typename T, int E, int TL, bool SPLIT, bool WAVE = false>
__device__ __forceinline__ void fetch_partner(const cplx<T> (&v)[E], cplx<T> (&p)[E],
                                              int t, T *re, T *im) {
  if constexpr (SPLIT) {
#pragma unroll
    for (int e = 0; e < E; ++e) re[row_slot<TL>(t, e)] = v[e].x;
    lds_sync<WAVE>();
#pragma unroll
    for (int e = 0; e < E; ++e) p[e].x = re[row_mirror_slot<E, TL>(t, e)];
    lds_sync<WAVE>();
#pragma unroll
    for (int e = 0; e < E; ++e) re[row_slot<TL>(t, e)] = v[e].y;
    lds_sync<WAVE>();
#pragma unroll
    for (int e = 0; e < E; ++e) p[e].y = re[row_mirror_slot<E, TL>(t, e)];
    lds_sync<WAVE>();
  } else {
    row_fetch_partner<T, E, TL, WAVE>(v, p, t, re, im);
  }
}

template <typename T> struct TStepArgs {
  Nonlin<T> nl;     // nonlinearity (plain or mean-split)
  int zero_k0;      // mean-split step: keep the k = 0 mode of the deviation at 0
  int mode;         // k_tz kind (ZMode)
  int nfields;      // k_ty_fwd: 2 (A and B) or 1 (A only)
  TungOp top;       // Tungsten operator scalars
  TungNl<T> tnl;    // Tungsten nonlinearity
  T *mf_out;        // k_tx_tung: optional copy of psi_MF, or null
  const T *psi_in;  // k_tx_fwd / k_tx_r2c / k_tx_tung input
  T *psi_out;       // k_tx_inv output
  cplx<T> *A1, *B1; // T1 buffers (Psi, N)
  cplx<T> *A2, *B2; // T2 buffers (Psi, N)
  const cplx<T> *tw_x;  // nx-point line twiddles (forward x)
  const cplx<T> *tw_xm; // nx/2-point line twiddles (inverse x)
  const cplx<T> *wn;    // exp(-2 pi i k/nx), k = 0..nx/2
  const cplx<T> *tw_y, *tw_z;
  PfcOp op;
  int nx, ny, nz, W;
};

// Kernels on a plan with double butterflies (Plan::XP) also evaluate the
// pointwise arithmetic that would otherwise carry a rounded constant in
// double -- the mean-split nonlinearity (from the unrounded 3m, 3m^2), the
// ETD combine (opL, opN unrounded) and the C2R merge (exp(2 pi i k/nx) from
// a double table) -- so that no FP32-rounded coefficient acts on the field
// step after step; each result is rounded once.
template <typename Plan, typename T>
__device__ __forceinline__ T nl_value(const Nonlin<T> &nl, T u) {
  if constexpr (Plan::XP)
    return (T)nl.exact((double)u);
  else
    return nl(u);
}

// ----------------------------------------------------------------------------
// psi -> T1: G consecutive rows y0..y0+G-1 of plane z per block.
// ----------------------------------------------------------------------------
template <typename T, typename Plan, int G, bool SPLIT, bool TILE>
__global__ void __launch_bounds__(G *Plan::TL) k_tx_fwd(TStepArgs<T> a) {
  constexpr int TL = Plan::TL, E = Plan::E, N = Plan::N, RP = row_padded(N);
  constexpr int PL = t_plane<Plan, G>();
  __shared__ T lds[(SPLIT ? 1 : 2) * PL];
  const int t = threadIdx.x % TL, g = threadIdx.x / TL;
  const int ngy = (a.ny + G - 1) / G;
  const unsigned bid = xcd_block<T, G>(blockIdx.x, gridDim.x);
  const int z = (int)bid / ngy, y0 = ((int)bid % ngy) * G;
  const int y = y0 + g;
  const bool active = y < a.ny;
  const T *__restrict__ row = a.psi_in + ((long)z * a.ny + (active ? y : y0)) * N;

  cplx<T> v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const T u = row[t + TL * e];
    v[e] = cplx<T>{u, nl_value<Plan>(a.nl, u)};
  }
  T *re = lds + g * RP;
  T *im = lds + (SPLIT ? 0 : PL) + g * RP;
  const LdsLine<T> ll{lds, g * RP, 1, PL};
  LineFft<T, Plan, -1, SPLIT, true, row_wave<Plan>()>::run(v, t, ll, a.tw_x);
  cplx<T> pz[E];
  fetch_partner<T, E, TL, SPLIT, row_wave<Plan>()>(v, pz, t, re, im);
  // split Z into Psi (v) and N (pz) for k = t + TL*e, e < E/2
#pragma unroll
  for (int e = 0; e <= E / 2 && e < E; ++e) {
    const cplx<T> zk = v[e], zc = c_conj(pz[e]);
    const cplx<T> s = c_add(zk, zc), d = c_sub(zk, zc);
    v[e] = c_scale(s, (T)0.5);
    pz[e] = cplx<T>{(T)0.5 * d.y, (T)-0.5 * d.x};
  }
  // k = N/2 (owned by t = 0 as element E/2): written directly
  if (active && t == 0) {
    const long o = ((long)z * a.W + N / 2) * a.ny + y;
    a.A1[o] = v[E / 2];
    a.B1[o] = pz[E / 2];
  }
  if constexpr (!TILE) {
    // direct strided stores: the G lines of the block fill each chunk
    if (active) {
#pragma unroll
      for (int r = 0; r < E / 2; ++r) {
        const long o = ((long)z * a.W + TL * r + t) * a.ny + y;
        a.A1[o] = v[r];
        a.B1[o] = pz[r];
      }
    }
    return;
  }
  // transposed store of kx = TL*r + j, j < TL, r < E/2, through a TL x G tile
  __syncthreads(); // every line's transform is done with the LDS
  T *tr = lds, *ti = lds + TL * tile_pitch<G, T>();
#pragma unroll
  for (int f = 0; f < 2; ++f) {
    cplx<T> *__restrict__ out = f ? a.B1 : a.A1;
#pragma unroll
    for (int r = 0; r < E / 2; ++r) {
      const cplx<T> w = f ? pz[r] : v[r];
      tr[tile_idx<G, T>(t, g)] = w.x;
      ti[tile_idx<G, T>(t, g)] = w.y;
      __syncthreads();
      const int j = threadIdx.x / G, gg = threadIdx.x % G;
      if (y0 + gg < a.ny) {
        const long o = ((long)z * a.W + TL * r + j) * a.ny + y0 + gg;
        out[o] = cplx<T>{tr[tile_idx<G, T>(j, gg)], ti[tile_idx<G, T>(j, gg)]};
      }
      __syncthreads();
    }
  }
}

// ----------------------------------------------------------------------------
// T1 -> T2 forward y: block = (field, kx, z0..z0+G-1).
// ----------------------------------------------------------------------------
template <typename T, typename Plan, int G, bool SPLIT, bool TILE>
__global__ void __launch_bounds__(G *Plan::TL) k_ty_fwd(TStepArgs<T> a) {
  constexpr int TL = Plan::TL, E = Plan::E, N = Plan::N, RP = row_padded(N);
  constexpr int PL = t_plane<Plan, G>();
  __shared__ T lds[(SPLIT ? 1 : 2) * PL];
  const int t = threadIdx.x % TL, g = threadIdx.x / TL;
  const int ngz = (a.nz + G - 1) / G;
  const unsigned bid = xcd_block<T, G>(blockIdx.x, gridDim.x);
  const int kx = (int)bid / ngz, z0 = ((int)bid % ngz) * G;
  const int z = z0 + g;
  const bool active = z < a.nz;
  const cplx<T> *__restrict__ in = blockIdx.y ? a.B1 : a.A1;
  cplx<T> *__restrict__ out = blockIdx.y ? a.B2 : a.A2;
  const cplx<T> *line = in + ((long)(active ? z : z0) * a.W + kx) * N;

  cplx<T> v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = line[t + TL * e];
  const LdsLine<T> ll{lds, g * RP, 1, PL};
  LineFft<T, Plan, -1, SPLIT, true, row_wave<Plan>()>::run(v, t, ll, a.tw_y);
  if constexpr (!TILE) {
    if (active) {
#pragma unroll
      for (int r = 0; r < E; ++r) out[((long)kx * N + TL * r + t) * a.nz + z] = v[r];
    }
    return;
  }
  // transposed store: T2[(kx*ny + ky)*nz + z], ky = TL*r + j
  __syncthreads(); // every line's transform is done with the LDS
  T *tr = lds, *ti = lds + TL * tile_pitch<G, T>();
#pragma unroll
  for (int r = 0; r < E; ++r) {
    tr[tile_idx<G, T>(t, g)] = v[r].x;
    ti[tile_idx<G, T>(t, g)] = v[r].y;
    __syncthreads();
    const int j = threadIdx.x / G, gg = threadIdx.x % G;
    if (z0 + gg < a.nz) {
      const long o = ((long)kx * N + TL * r + j) * a.nz + z0 + gg;
      out[o] = cplx<T>{tr[tile_idx<G, T>(j, gg)], ti[tile_idx<G, T>(j, gg)]};
    }
    __syncthreads();
  }
}

// ----------------------------------------------------------------------------
// T2 z pass, in place: G consecutive z lines (kx*ny + ky) per block.
//
// NT = true (here and in k_tx_inv, k_tx_fwd_w, k_ty_fwd_w, k_ty_inv_w) marks
// the loads and stores of lines non-temporal (ld_line / st_line): every such
// value is read or written once per pass, and the hint keeps the stream from
// displacing the twiddles and the lines other blocks still share.
// ----------------------------------------------------------------------------
template <bool NT, typename T> __device__ __forceinline__ cplx<T> ld_line(const cplx<T> *p) {
  if constexpr (NT)
    return cplx<T>{__builtin_nontemporal_load(&p->x), __builtin_nontemporal_load(&p->y)};
  else
    return *p;
}
template <bool NT, typename T> __device__ __forceinline__ T ld_val(const T *p) {
  if constexpr (NT)
    return __builtin_nontemporal_load(p);
  else
    return *p;
}
template <bool NT, typename T> __device__ __forceinline__ void st_line(cplx<T> *p, cplx<T> v) {
  if constexpr (NT) {
    __builtin_nontemporal_store(v.x, &p->x);
    __builtin_nontemporal_store(v.y, &p->y);
  } else {
    *p = v;
  }
}

// Block-wide load of a ROWS x G transpose tile (sr/si at tile_idx<G, T>(row,
// line)) from HBM, NTH threads, element i = row*G + line.  Every load of a
// thread is issued before the first LDS write, so a thread keeps
// ceil(ROWS*G / NTH) HBM requests in flight instead of one per loop trip:
// the addresses of lines >= nvalid and of the trip past the tile's end are
// clamped to element (0, 0) of the tile (always valid) and their values
// dropped, which keeps the issue loop free of branches.  (The stores of a
// tile stay a plain loop: a store does not stall the thread, and holding
// the values of several trips costs registers the wide kernels do not have.)
template <int ROWS, int G, int NTH, bool NT, typename T, typename Src>
__device__ __forceinline__ void tile_load(T *sr, T *si, int nvalid, Src src) {
  constexpr int TOT = ROWS * G, IT = (TOT + NTH - 1) / NTH;
  cplx<T> w[IT];
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int i = (int)threadIdx.x + it * NTH;
    const int r = i / G, gg = i % G;
    const bool ok = (TOT % NTH == 0 || i < TOT) && gg < nvalid;
    const cplx<T> v = ld_line<NT>(ok ? src(r, gg) : src(0, 0));
    w[it] = ok ? v : cplx<T>{(T)0, (T)0};
  }
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int i = (int)threadIdx.x + it * NTH;
    if (TOT % NTH == 0 || i < TOT) {
      sr[tile_idx<G, T>(i / G, i % G)] = w[it].x;
      si[tile_idx<G, T>(i / G, i % G)] = w[it].y;
    }
  }
}

// Block-wide store of a ROWS x G transpose tile (sr/si at tile_idx<G, T>(row,
// line)) to HBM, NTH threads, element i = row*G + line: every LDS read of a
// thread is issued before its first store, so the stores leave back to back
// instead of one LDS round trip apart.  Lines >= nvalid are not stored.  It
// holds ceil(ROWS*G / NTH) values per thread, which the wide forward y pass
// has room for with 8-byte planes (122 -> 126 VGPRs, four waves per SIMD
// either way) but not with 4-byte ones (61 -> 66: one of its 16-wave blocks
// per CU instead of two), nor do the x passes (their stores follow the
// transforms with more values live).
template <int ROWS, int G, int NTH, bool NT, typename T, typename Dst>
__device__ __forceinline__ void tile_store(const T *sr, const T *si, int nvalid, Dst dst) {
  constexpr int TOT = ROWS * G, IT = (TOT + NTH - 1) / NTH;
  cplx<T> w[IT];
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int i = (int)threadIdx.x + it * NTH;
    if (TOT % NTH == 0 || i < TOT)
      w[it] = cplx<T>{sr[tile_idx<G, T>(i / G, i % G)], si[tile_idx<G, T>(i / G, i % G)]};
  }
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int i = (int)threadIdx.x + it * NTH;
    if ((TOT % NTH == 0 || i < TOT) && i % G < nvalid) st_line<NT>(dst(i / G, i % G), w[it]);
  }
}

template <typename T, typename Plan, int G, bool SPLIT, int MODE, bool PRE = false,
          bool NT = false>
__global__ void __launch_bounds__(G *Plan::TL) k_tz(TStepArgs<T> a) {
  __shared__ T lds[(SPLIT ? 1 : 2) * t_plane<Plan, G>()];
  constexpr int TL = Plan::TL, E = Plan::E, N = Plan::N, RP = row_padded(N);
  constexpr int PL = t_plane<Plan, G>();
  const int t = threadIdx.x % TL, g = threadIdx.x / TL;
  const long nlines = (long)a.W * a.ny;
  const long l0 = (long)blockIdx.x * G;
  const long li = l0 + g;
  const bool active = li < nlines;
  const long l = active ? li : l0;
  const int kx = (int)(l / a.ny), ky = (int)(l % a.ny);
  const LdsLine<T> ll{lds, g * RP, 1, PL};
  cplx<T> *__restrict__ la = a.A2 + l * N;
  cplx<T> *__restrict__ lb = a.B2 + l * N;
  const double kxy = pfc_kxy(a.op, kx, ky);

  cplx<T> acc[E];
  if constexpr (MODE == ZM_TMF) {
    // A = Fz(A) (kept for the combine); B = Fz^-1(fMF * A)
#pragma unroll
    for (int e = 0; e < E; ++e) acc[e] = ld_line<NT>(la + (t + TL * e));
    LineFft<T, Plan, -1, SPLIT, true, row_wave<Plan>()>::run(acc, t, ll, a.tw_z);
    if (active) {
#pragma unroll
      for (int e = 0; e < E; ++e) st_line<NT>(la + (t + TL * e), acc[e]);
    }
#pragma unroll
    for (int e = 0; e < E; ++e) {
      double f, unused;
      z_mode_ops<ZM_TMF>(a.op, a.top, kxy, pfc_kc(t + TL * e, a.op.nz, a.op.fz), f,
                         unused);
      acc[e] = c_scale(acc[e], (T)f);
    }
  } else {
    // PRE = false: multipliers first, while nothing else is live (exp and
    // division need many temporaries), then the lines are loaded.  PRE =
    // true: both lines are requested first and the multipliers computed
    // while the loads are in flight.
    // XP plans keep the multipliers unrounded (double) and combine in double
    using M = std::conditional_t<Plan::XP, double, T>;
    M opl[E], opn[E];
    cplx<T> v[E];
    auto multipliers = [&] {
#pragma unroll
      for (int e = 0; e < E; ++e) {
        double ma, mb;
        z_mode_ops<MODE>(a.op, a.top, kxy, pfc_kc(t + TL * e, a.op.nz, a.op.fz), ma, mb);
        opl[e] = (M)ma;
        opn[e] = (M)mb;
      }
    };
    if constexpr (PRE) {
#pragma unroll
      for (int e = 0; e < E; ++e) {
        acc[e] = ld_line<NT>(la + (t + TL * e));
        v[e] = ld_line<NT>(lb + (t + TL * e));
      }
      multipliers();
    } else {
      multipliers();
#pragma unroll
      for (int e = 0; e < E; ++e) {
        if constexpr (MODE == ZM_PFC) acc[e] = ld_line<NT>(la + (t + TL * e));
        v[e] = ld_line<NT>(lb + (t + TL * e));
      }
    }
    if constexpr (MODE == ZM_PFC) {
      // A = Fz^-1(opL*Fz(A) + opN*Fz(B))
      LineFft<T, Plan, -1, SPLIT, true, row_wave<Plan>()>::run(acc, t, ll, a.tw_z);
      if constexpr (!Plan::XP) {
#pragma unroll
        for (int e = 0; e < E; ++e) acc[e] = c_scale(acc[e], opl[e]);
      }
      LineFft<T, Plan, -1, SPLIT, true, row_wave<Plan>()>::run(v, t, ll, a.tw_z);
#pragma unroll
      for (int e = 0; e < E; ++e) {
        if constexpr (Plan::XP)
          acc[e] = cplx<T>{(T)((double)acc[e].x * opl[e] + (double)v[e].x * opn[e]),
                           (T)((double)acc[e].y * opl[e] + (double)v[e].y * opn[e])};
        else
          acc[e] = c_add(acc[e], c_scale(v[e], opn[e]));
      }
      if (a.zero_k0 && l == 0 && t == 0) acc[0] = cplx<T>{0, 0};
    } else {
      // A = Fz^-1(expL*A + nW*Fz(B)), A already a full spectrum
      LineFft<T, Plan, -1, SPLIT, true, row_wave<Plan>()>::run(v, t, ll, a.tw_z);
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const cplx<T> h = PRE ? acc[e] : ld_line<NT>(la + (t + TL * e));
        acc[e] = c_add(c_scale(h, (T)opl[e]), c_scale(v[e], (T)opn[e]));
      }
    }
  }
  LineFft<T, Plan, +1, SPLIT, true, row_wave<Plan>()>::run(acc, t, ll, a.tw_z);
  if (active) {
    cplx<T> *__restrict__ out = MODE == ZM_TMF ? lb : la;
#pragma unroll
    for (int e = 0; e < E; ++e) st_line<NT>(out + (t + TL * e), acc[e]);
  }
}

// ----------------------------------------------------------------------------
// T2 z pass as k_tz (PRE = true ordering, non-temporal lines), with the
// multipliers of each line tabulated in LDS.  kz enters every mode's
// multipliers only as kz*kz, and pfc_kc(k) = -pfc_kc(nz - k) exactly, so a
// line has nz/2 + 1 distinct pairs: the block evaluates those once (k_tz
// evaluates all nz), while its loads are in flight, and each transform's
// output reads its pair back from slot k or nz - k.  No multiplier stays in
// registers across a transform.  The values and the arithmetic that applies
// them are k_tz's, so both give the same bits.
//
// PAIR (G = 2): the block's lines are (kx, ky) and (kx, ny - ky), whose
// kx*kx + ky*ky agree bit for bit by the same argument in y, so one table
// serves both and the two lines' threads fill it together (half the
// evaluations again).  Block 0 of each kx holds the lines that are their own
// mirror, ky = 0 and (ny even) ky = ny/2, with a table each; for odd ny its
// second line is idle.  Grid: W * ((ny + 1) / 2) blocks.
// ----------------------------------------------------------------------------
template <typename T, typename Plan, int G> constexpr int tz_table_bytes() {
  return G * 2 * (Plan::N / 2 + 1) * (int)sizeof(std::conditional_t<Plan::XP, double, T>);
}

template <typename T, typename Plan, int G, bool SPLIT, int MODE, bool PAIR = false>
__global__ void __launch_bounds__(G *Plan::TL) k_tz_mt(TStepArgs<T> a) {
  static_assert(!PAIR || G == 2, "a paired block holds two lines");
  constexpr int TL = Plan::TL, E = Plan::E, N = Plan::N, RP = row_padded(N);
  constexpr int PL = t_plane<Plan, G>(), NH = N / 2 + 1;
  using M = std::conditional_t<Plan::XP, double, T>;
  __shared__ T lds[(SPLIT ? 1 : 2) * PL];
  __shared__ M tab[G * 2 * NH];
  const int t = threadIdx.x % TL, g = threadIdx.x / TL;
  long l;
  bool active, shared = false;
  if constexpr (PAIR) {
    const int nb = (a.ny + 1) / 2; // blocks per kx
    const int kxb = (int)(blockIdx.x / nb), j = (int)(blockIdx.x % nb);
    shared = j != 0;
    const int kyp = j != 0 ? (g == 0 ? j : a.ny - j) : (g == 0 ? 0 : a.ny / 2);
    active = j != 0 || g == 0 || a.ny % 2 == 0;
    l = (long)kxb * a.ny + (active ? kyp : 0);
  } else {
    const long nlines = (long)a.W * a.ny;
    const long l0 = (long)blockIdx.x * G;
    const long li = l0 + g;
    active = li < nlines;
    l = active ? li : l0;
  }
  const int kx = (int)(l / a.ny), ky = (int)(l % a.ny);
  const LdsLine<T> ll{lds, g * RP, 1, PL};
  cplx<T> *__restrict__ la = a.A2 + l * N;
  cplx<T> *__restrict__ lb = a.B2 + l * N;
  M *__restrict__ ta = tab + (shared ? 0 : g) * 2 * NH;
  M *__restrict__ tb = ta + NH;
  auto slot = [&](int e) {
    const int k = t + TL * e;
    return k <= N / 2 ? k : N - k;
  };

  cplx<T> acc[E], v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    acc[e] = ld_line<true>(la + (t + TL * e));
    if constexpr (MODE != ZM_TMF) v[e] = ld_line<true>(lb + (t + TL * e));
  }
  {
    const double kxy = pfc_kxy(a.op, kx, ky);
    const int k0 = shared ? (int)threadIdx.x : t, dk = shared ? G * TL : TL;
    for (int k = k0; k < NH; k += dk) {
      double ma, mb;
      z_mode_ops<MODE>(a.op, a.top, kxy, pfc_kc(k, a.op.nz, a.op.fz), ma, mb);
      ta[k] = (M)ma;
      if constexpr (MODE != ZM_TMF) tb[k] = (M)mb;
    }
  }
  __syncthreads(); // the table is complete
  if constexpr (MODE == ZM_TMF) {
    // A = Fz(A) (kept for the combine); B = Fz^-1(fMF * A)
    LineFft<T, Plan, -1, SPLIT, true, row_wave<Plan>()>::run(acc, t, ll, a.tw_z);
    if (active) {
#pragma unroll
      for (int e = 0; e < E; ++e) st_line<true>(la + (t + TL * e), acc[e]);
    }
#pragma unroll
    for (int e = 0; e < E; ++e) acc[e] = c_scale(acc[e], (T)ta[slot(e)]);
  } else if constexpr (MODE == ZM_PFC) {
    // A = Fz^-1(opL*Fz(A) + opN*Fz(B))
    LineFft<T, Plan, -1, SPLIT, true, row_wave<Plan>()>::run(acc, t, ll, a.tw_z);
    if constexpr (!Plan::XP) {
#pragma unroll
      for (int e = 0; e < E; ++e) acc[e] = c_scale(acc[e], ta[slot(e)]);
    }
    LineFft<T, Plan, -1, SPLIT, true, row_wave<Plan>()>::run(v, t, ll, a.tw_z);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const M opn = tb[slot(e)];
      if constexpr (Plan::XP) {
        const M opl = ta[slot(e)];
        acc[e] = cplx<T>{(T)((double)acc[e].x * opl + (double)v[e].x * opn),
                         (T)((double)acc[e].y * opl + (double)v[e].y * opn)};
      } else {
        acc[e] = c_add(acc[e], c_scale(v[e], opn));
      }
    }
    if (a.zero_k0 && l == 0 && t == 0) acc[0] = cplx<T>{0, 0};
  } else {
    // A = Fz^-1(expL*A + nW*Fz(B)), A already a full spectrum
    LineFft<T, Plan, -1, SPLIT, true, row_wave<Plan>()>::run(v, t, ll, a.tw_z);
#pragma unroll
    for (int e = 0; e < E; ++e)
      acc[e] = c_add(c_scale(acc[e], (T)ta[slot(e)]), c_scale(v[e], (T)tb[slot(e)]));
  }
  LineFft<T, Plan, +1, SPLIT, true, row_wave<Plan>()>::run(acc, t, ll, a.tw_z);
  if (active) {
    cplx<T> *__restrict__ out = MODE == ZM_TMF ? lb : la;
#pragma unroll
    for (int e = 0; e < E; ++e) st_line<true>(out + (t + TL * e), acc[e]);
  }
}


// ----------------------------------------------------------------------------
// T2 -> T1 inverse y: block = (kx, z0..z0+G-1); transposed load.
// ----------------------------------------------------------------------------
template <typename T, typename Plan, int G, bool SPLIT, bool TILE>
__global__ void __launch_bounds__(G *Plan::TL) k_ty_inv(TStepArgs<T> a) {
  constexpr int TL = Plan::TL, E = Plan::E, N = Plan::N, RP = row_padded(N);
  constexpr int PL = t_plane<Plan, G>();
  __shared__ T lds[(SPLIT ? 1 : 2) * PL];
  const int t = threadIdx.x % TL, g = threadIdx.x / TL;
  const int ngz = (a.nz + G - 1) / G;
  const unsigned bid = xcd_block<T, G>(blockIdx.x, gridDim.x);
  const int kx = (int)bid / ngz, z0 = ((int)bid % ngz) * G;
  const int z = z0 + g;

  cplx<T> v[E];
  if constexpr (!TILE) {
    const int zz = z < a.nz ? z : z0;
#pragma unroll
    for (int r = 0; r < E; ++r) v[r] = a.A2[((long)kx * N + TL * r + t) * a.nz + zz];
  }
  T *tr = lds, *ti = lds + TL * tile_pitch<G, T>();
#pragma unroll
  for (int r = 0; r < (TILE ? E : 0); ++r) {
    const int j = threadIdx.x / G, gg = threadIdx.x % G;
    cplx<T> w{0, 0};
    if (z0 + gg < a.nz) w = a.A2[((long)kx * N + TL * r + j) * a.nz + z0 + gg];
    tr[tile_idx<G, T>(j, gg)] = w.x;
    ti[tile_idx<G, T>(j, gg)] = w.y;
    __syncthreads();
    v[r] = cplx<T>{tr[tile_idx<G, T>(t, g)], ti[tile_idx<G, T>(t, g)]};
    __syncthreads();
  }
  const LdsLine<T> ll{lds, g * RP, 1, PL};
  LineFft<T, Plan, +1, SPLIT, true, row_wave<Plan>()>::run(v, t, ll, a.tw_y);
  if (z < a.nz) {
    cplx<T> *__restrict__ line = a.A1 + ((long)z * a.W + kx) * N;
#pragma unroll
    for (int e = 0; e < E; ++e) line[t + TL * e] = v[e];
  }
}

// ----------------------------------------------------------------------------
// T1 -> psi inverse x (C2R): block = (z, y0..y0+G-1); the G rows' half
// spectra (W values each) are staged in LDS, then each line runs the
// nx/2-point inverse.  Imaginary parts of X[0] and X[nx/2] are ignored.
// Plan is the M = nx/2 point plan.
// ----------------------------------------------------------------------------
template <typename T, typename Plan, int G, bool TILE> constexpr int tx_inv_lds() {
  constexpr int stage = TILE ? 2 * (Plan::N + 1) * tile_pitch<G, T>() : 0;
  constexpr int fft = 2 * t_plane<Plan, G>();
  return stage > fft ? stage : fft;
}

template <typename T, typename Plan, int G, bool TILE, bool NT = false>
__global__ void __launch_bounds__(G *Plan::TL) k_tx_inv(TStepArgs<T> a) {
  __shared__ T lds[tx_inv_lds<T, Plan, G, TILE>()];
  constexpr int TL = Plan::TL, E = Plan::E, M = Plan::N, RP = row_padded(M);
  constexpr int PL = t_plane<Plan, G>();
  constexpr int W = M + 1;
  const int t = threadIdx.x % TL, g = threadIdx.x / TL;
  const int ngy = (a.ny + G - 1) / G;
  const unsigned bid = xcd_block<T, G>(blockIdx.x, gridDim.x);
  const int z = (int)bid / ngy, y0 = ((int)bid % ngy) * G;
  const int y = y0 + g;

  cplx<T> v[E];
  if constexpr (!TILE) {
    // direct strided loads of X[k] and X[M-k] of this line
    const int yy = y < a.ny ? y : y0;
    const cplx<T> *__restrict__ col = a.A1 + (long)z * a.W * a.ny + yy;
    using C = std::conditional_t<Plan::XP, double, T>; // merge arithmetic
    const cplx<C> *wn = reinterpret_cast<const cplx<C> *>(a.wn);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = t + TL * e;
      const cplx<T> ca = col[(long)k * a.ny], cb = col[(long)(M - k) * a.ny];
      cplx<C> xa{(C)ca.x, (C)ca.y};
      cplx<C> xb{(C)cb.x, -(C)cb.y};
      if (k == 0) {
        xa.y = (C)0;
        xb.y = (C)0;
      }
      const cplx<C> ev = c_add(xa, xb);
      const cplx<C> od = c_mulc(c_sub(xa, xb), wn[k]);
      v[e] = cplx<T>{(T)(ev.x - od.y), (T)(ev.y + od.x)};
    }
  } else {
  // stage: sr/si[kx*(G+1) + line]
  T *sr = lds, *si = lds + W * tile_pitch<G, T>();
  const cplx<T> *__restrict__ src = a.A1 + (long)z * a.W * a.ny + y0;
  tile_load<W, G, G * TL, NT>(sr, si, a.ny - y0,
                              [&](int kx, int gg) { return src + ((long)kx * a.ny + gg); });
  __syncthreads();
  using C = std::conditional_t<Plan::XP, double, T>; // merge arithmetic
  const cplx<C> *wn = reinterpret_cast<const cplx<C> *>(a.wn);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int k = t + TL * e;
    cplx<C> xa{(C)sr[tile_idx<G, T>(k, g)], (C)si[tile_idx<G, T>(k, g)]};
    cplx<C> xb{(C)sr[tile_idx<G, T>(M - k, g)], -(C)si[tile_idx<G, T>(M - k, g)]};
    if (k == 0) {
      xa.y = (C)0;
      xb.y = (C)0;
    }
    const cplx<C> ev = c_add(xa, xb);
    const cplx<C> od = c_mulc(c_sub(xa, xb), wn[k]);
    v[e] = cplx<T>{(T)(ev.x - od.y), (T)(ev.y + od.x)};
  }
  __syncthreads(); // the staging area becomes the FFT exchange area
  }
  const LdsLine<T> ll{lds, g * RP, 1, PL};
  LineFft<T, Plan, +1, false, true, row_wave<Plan>()>::run(v, t, ll, a.tw_xm);
  if (y < a.ny) {
    cplx<T> *__restrict__ out =
        reinterpret_cast<cplx<T> *>(a.psi_out + ((long)z * a.ny + y) * (2L * M));
#pragma unroll
    for (int e = 0; e < E; ++e) st_line<NT>(out + (t + TL * e), v[e]);
  }
}


// ----------------------------------------------------------------------------
// Half-spectrum rows of G consecutive y rows of plane z, transposed store into
// T1 (out[(z*W + kx)*ny + y]): v[r] holds X[TL*r + t] of row y0 + g, and
// xm = X[M] (thread t == 0).  Through a TL x G LDS tile when TILE.
// ----------------------------------------------------------------------------
template <typename T, int E, int TL, int G, bool TILE>
__device__ __forceinline__ void t1_store_rows(const cplx<T> (&v)[E], cplx<T> xm,
                                              cplx<T> *__restrict__ out, T *lds,
                                              int t, int g, int z, int y0,
                                              int ny, int W) {
  constexpr int M = E * TL;
  const int y = y0 + g;
  if (y < ny && t == 0) out[((long)z * W + M) * ny + y] = xm;
  if constexpr (!TILE) {
    if (y < ny) {
#pragma unroll
      for (int r = 0; r < E; ++r) out[((long)z * W + TL * r + t) * ny + y] = v[r];
    }
  } else {
    __syncthreads(); // every line's transform is done with the LDS
    T *tr = lds, *ti = lds + TL * tile_pitch<G, T>();
#pragma unroll
    for (int r = 0; r < E; ++r) {
      tr[tile_idx<G, T>(t, g)] = v[r].x;
      ti[tile_idx<G, T>(t, g)] = v[r].y;
      __syncthreads();
      const int j = threadIdx.x / G, gg = threadIdx.x % G;
      if (y0 + gg < ny)
        out[((long)z * W + TL * r + j) * ny + y0 + gg] =
            cplx<T>{tr[tile_idx<G, T>(j, gg)], ti[tile_idx<G, T>(j, gg)]};
      __syncthreads();
    }
  }
}

// The same store through one (M+1) x G LDS tile: two block-wide syncs.
template <typename T, int E, int TL, int G>
__device__ __forceinline__ void t1_store_rows_wide(const cplx<T> (&v)[E], cplx<T> xm,
                                                   cplx<T> *__restrict__ out, T *lds, int t,
                                                   int g, int z, int y0, int ny, int W) {
  constexpr int M = E * TL, WR = M + 1;
  T *sr = lds, *si = lds + WR * tile_pitch<G, T>();
  __syncthreads(); // previous users of the LDS are done
#pragma unroll
  for (int e = 0; e < E; ++e) {
    sr[tile_idx<G, T>(t + TL * e, g)] = v[e].x;
    si[tile_idx<G, T>(t + TL * e, g)] = v[e].y;
  }
  if (t == 0) {
    sr[tile_idx<G, T>(M, g)] = xm.x;
    si[tile_idx<G, T>(M, g)] = xm.y;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < WR * G; i += G * TL) {
    const int kx = i / G, gg = i % G;
    if (y0 + gg < ny)
      out[((long)z * W + kx) * ny + y0 + gg] =
          cplx<T>{sr[tile_idx<G, T>(kx, gg)], si[tile_idx<G, T>(kx, gg)]};
  }
}

// Forward half-length transform of the G rows held as v (z[m] = x[2m] +
// i*x[2m+1]) and the R2C split; on return v[e] = X[t + TL*e] and xm = X[M]
// (valid on t == 0).
template <typename T, typename Plan, int G>
__device__ __forceinline__ void row_r2c(cplx<T> (&v)[Plan::E], cplx<T> &xm, T *lds,
                                        int t, int g, const cplx<T> *tw,
                                        const cplx<T> *wn) {
  constexpr int TL = Plan::TL, E = Plan::E, M = Plan::N, RP = row_padded(M);
  constexpr int PL = t_plane<Plan, G>();
  const LdsLine<T> ll{lds, g * RP, 1, PL};
  LineFft<T, Plan, -1, false, true, row_wave<Plan>()>::run(v, t, ll, tw);
  cplx<T> pz[E];
  row_fetch_partner<T, E, TL, row_wave<Plan>()>(v, pz, t, lds + g * RP, lds + PL + g * RP);
  xm = cplx<T>{v[0].x - v[0].y, (T)0};
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = r2c_split(v[e], pz[e], wn[t + TL * e]);
}

// ----------------------------------------------------------------------------
// Tungsten, psi -> T1 (A1): R2C of G consecutive rows of plane z with the
// nx/2-point plan, transposed store.
// ----------------------------------------------------------------------------
template <typename T, typename Plan, int G, bool TILE, bool WIDE = false>
__global__ void __launch_bounds__(G *Plan::TL) k_tx_r2c(TStepArgs<T> a) {
  constexpr int TL = Plan::TL, E = Plan::E, M = Plan::N;
  __shared__ T lds[WIDE ? tx_inv_lds<T, Plan, G, true>() : 2 * t_plane<Plan, G>()];
  const int t = threadIdx.x % TL, g = threadIdx.x / TL;
  const int ngy = (a.ny + G - 1) / G;
  const unsigned bid = xcd_block<T, G>(blockIdx.x, gridDim.x);
  const int z = (int)bid / ngy, y0 = ((int)bid % ngy) * G;
  const int y = y0 + g < a.ny ? y0 + g : y0;
  const cplx<T> *__restrict__ row =
      reinterpret_cast<const cplx<T> *>(a.psi_in + ((long)z * a.ny + y) * (2L * M));
  cplx<T> v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = row[t + TL * e];
  cplx<T> xm;
  row_r2c<T, Plan, G>(v, xm, lds, t, g, a.tw_xm, a.wn);
  if constexpr (WIDE)
    t1_store_rows_wide<T, E, TL, G>(v, xm, a.A1, lds, t, g, z, y0, a.ny, a.W);
  else
    t1_store_rows<T, E, TL, G, TILE>(v, xm, a.A1, lds, t, g, z, y0, a.ny, a.W);
}

// ----------------------------------------------------------------------------
// Tungsten fused x pass on T1 (B1, in place): for G consecutive rows of plane
// z, the mean-field half spectra (inverse y and z done, 1/(nx ny nz) folded
// into fMF) are loaded transposed, brought to real space (nx/2-point
// inverse), combined with the rows of psi into N(psi, psi_MF), transformed
// forward and stored transposed.  psi_MF goes to HBM only when mf_out is set.
// ----------------------------------------------------------------------------
template <typename T, typename Plan, int G, bool TILE, bool WIDE>
__device__ __forceinline__ void tx_tung_body(const TStepArgs<T> &a, T *lds) {
  constexpr int TL = Plan::TL, E = Plan::E, M = Plan::N, RP = row_padded(M);
  constexpr int PL = t_plane<Plan, G>();
  constexpr int W = M + 1;
  const int t = threadIdx.x % TL, g = threadIdx.x / TL;
  const int ngy = (a.ny + G - 1) / G;
  const unsigned bid = xcd_block<T, G>(blockIdx.x, gridDim.x);
  const int z = (int)bid / ngy, y0 = ((int)bid % ngy) * G;
  const int y = y0 + g;
  const int yy = y < a.ny ? y : y0;
  cplx<T> *__restrict__ spec = a.B1;

  cplx<T> v[E];
  if constexpr (!TILE) {
    const cplx<T> *__restrict__ col = spec + (long)z * a.W * a.ny + yy;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = t + TL * e;
      v[e] = c2r_merge(col[(long)k * a.ny], col[(long)(M - k) * a.ny], a.wn[k], k == 0);
    }
  } else {
    T *sr = lds, *si = lds + W * tile_pitch<G, T>();
    const cplx<T> *__restrict__ src = spec + (long)z * a.W * a.ny + y0;
    tile_load<W, G, G * TL, false>(sr, si, a.ny - y0,
                                   [&](int kx, int gg) { return src + ((long)kx * a.ny + gg); });
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = t + TL * e;
      const cplx<T> xa{sr[tile_idx<G, T>(k, g)], si[tile_idx<G, T>(k, g)]};
      const cplx<T> xb{sr[tile_idx<G, T>(M - k, g)], si[tile_idx<G, T>(M - k, g)]};
      v[e] = c2r_merge(xa, xb, a.wn[k], k == 0);
    }
    __syncthreads(); // the staging area becomes the FFT exchange area
  }
  const LdsLine<T> ll{lds, g * RP, 1, PL};
  LineFft<T, Plan, +1, false, true, row_wave<Plan>()>::run(v, t, ll, a.tw_xm);
  // v[e] = (psi_MF[2k], psi_MF[2k+1]), k = t + TL*e
  const long ro = ((long)z * a.ny + yy) * (2L * M);
  if (a.mf_out && y < a.ny) {
    cplx<T> *__restrict__ mo = reinterpret_cast<cplx<T> *>(a.mf_out + ro);
#pragma unroll
    for (int e = 0; e < E; ++e) mo[t + TL * e] = v[e];
  }
  const cplx<T> *__restrict__ pr = reinterpret_cast<const cplx<T> *>(a.psi_in + ro);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const cplx<T> u = pr[t + TL * e];
    v[e] = cplx<T>{a.tnl(u.x, v[e].x), a.tnl(u.y, v[e].y)};
  }
  cplx<T> xm;
  row_r2c<T, Plan, G>(v, xm, lds, t, g, a.tw_xm, a.wn);
  if constexpr (WIDE)
    t1_store_rows_wide<T, E, TL, G>(v, xm, spec, lds, t, g, z, y0, a.ny, a.W);
  else
    t1_store_rows<T, E, TL, G, TILE>(v, xm, spec, lds, t, g, z, y0, a.ny, a.W);
}

template <typename T, typename Plan, int G, bool TILE, bool WIDE = false>
__global__ void __launch_bounds__(G *Plan::TL) k_tx_tung(TStepArgs<T> a) {
  __shared__ T lds[tx_inv_lds<T, Plan, G, TILE>()];
  tx_tung_body<T, Plan, G, TILE, WIDE>(a, lds);
}
// The same pass with registers capped so that WPE waves fit on a SIMD: with
// the 2 x (nx/2 + 1) x G staging tile, G = 8 lines of an E = 8 plan
// fit two blocks on a CU only at <= 128 VGPRs, which the compiler reaches
// without spilling when asked to.
template <typename T, typename Plan, int G, bool TILE, bool WIDE, int WPE>
__global__ void __launch_bounds__(G *Plan::TL, WPE) k_tx_tung_o(TStepArgs<T> a) {
  __shared__ T lds[tx_inv_lds<T, Plan, G, TILE>()];
  tx_tung_body<T, Plan, G, TILE, WIDE>(a, lds);
}

// ----------------------------------------------------------------------------
// PFC forward x as two half-length R2Cs: psi -> T1 (A1: Psi, B1: N).  The G
// rows of plane z are read once as nx/2 complex pairs; the transform of psi
// and then that of N(psi) run on the nx/2-point plan, each followed by its
// transposed store.  Same bytes as k_tx_fwd, half the transform length.
// ----------------------------------------------------------------------------
template <typename T, typename Plan, int G, bool TILE>
__global__ void __launch_bounds__(G *Plan::TL) k_tx_fwd2(TStepArgs<T> a) {
  constexpr int TL = Plan::TL, E = Plan::E, M = Plan::N;
  __shared__ T lds[2 * t_plane<Plan, G>()];
  const int t = threadIdx.x % TL, g = threadIdx.x / TL;
  const int ngy = (a.ny + G - 1) / G;
  const unsigned bid = xcd_block<T, G>(blockIdx.x, gridDim.x);
  const int z = (int)bid / ngy, y0 = ((int)bid % ngy) * G;
  const int y = y0 + g < a.ny ? y0 + g : y0;
  const cplx<T> *__restrict__ row =
      reinterpret_cast<const cplx<T> *>(a.psi_in + ((long)z * a.ny + y) * (2L * M));
  cplx<T> u[E], v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    u[e] = row[t + TL * e];
    v[e] = u[e];
  }
  cplx<T> xm;
  row_r2c<T, Plan, G>(v, xm, lds, t, g, a.tw_xm, a.wn);
  t1_store_rows<T, E, TL, G, TILE>(v, xm, a.A1, lds, t, g, z, y0, a.ny, a.W);
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = cplx<T>{a.nl(u[e].x), a.nl(u[e].y)};
  row_r2c<T, Plan, G>(v, xm, lds, t, g, a.tw_xm, a.wn);
  t1_store_rows<T, E, TL, G, TILE>(v, xm, a.B1, lds, t, g, z, y0, a.ny, a.W);
}

// The forward x of G rows held as v[e] = psi + i*N(psi) at positions
// t + TL*e: N-point transform, split into the two half spectra, and the wide
// transposed store of each into T1 (A1: Psi, B1: N) through one W x G tile.
template <typename T, typename Plan, int G, bool SPLIT, bool NT = false>
__device__ __forceinline__ void fwd_split_store_wide(cplx<T> (&v)[Plan::E], T *lds, int t, int g,
                                                     int z, int y0, const TStepArgs<T> &a) {
  constexpr int TL = Plan::TL, E = Plan::E, N = Plan::N, RP = row_padded(N);
  constexpr int PL = t_plane<Plan, G>();
  constexpr int W = N / 2 + 1;
  T *re = lds + g * RP;
  T *im = lds + (SPLIT ? 0 : PL) + g * RP;
  const LdsLine<T> ll{lds, g * RP, 1, PL};
  LineFft<T, Plan, -1, SPLIT, true, row_wave<Plan>()>::run(v, t, ll, a.tw_x);
  cplx<T> pz[E];
  fetch_partner<T, E, TL, SPLIT, row_wave<Plan>()>(v, pz, t, re, im);
#pragma unroll
  for (int e = 0; e <= E / 2 && e < E; ++e) {
    const cplx<T> zk = v[e], zc = c_conj(pz[e]);
    const cplx<T> sum = c_add(zk, zc), d = c_sub(zk, zc);
    v[e] = c_scale(sum, (T)0.5);
    pz[e] = cplx<T>{(T)0.5 * d.y, (T)-0.5 * d.x};
  }
  T *sr = lds, *si = lds + W * tile_pitch<G, T>();
#pragma unroll
  for (int f = 0; f < 2; ++f) {
    __syncthreads(); // the previous users of the LDS are done
#pragma unroll
    for (int e = 0; e < E / 2; ++e) {
      const cplx<T> w = f ? pz[e] : v[e];
      sr[tile_idx<G, T>(t + TL * e, g)] = w.x;
      si[tile_idx<G, T>(t + TL * e, g)] = w.y;
    }
    if (t == 0) { // k = N/2, element E/2 of thread 0
      const cplx<T> w = f ? pz[E / 2] : v[E / 2];
      sr[tile_idx<G, T>(N / 2, g)] = w.x;
      si[tile_idx<G, T>(N / 2, g)] = w.y;
    }
    __syncthreads();
    cplx<T> *__restrict__ out = f ? a.B1 : a.A1;
    for (int i = threadIdx.x; i < W * G; i += G * TL) {
      const int kx = i / G, gg = i % G;
      if (y0 + gg < a.ny)
        st_line<NT>(out + (((long)z * a.W + kx) * a.ny + y0 + gg),
                    cplx<T>{sr[tile_idx<G, T>(kx, gg)], si[tile_idx<G, T>(kx, gg)]});
    }
  }
}

// ----------------------------------------------------------------------------
// k_tx_fwd with a wide transposed store: per field, all W half-spectrum
// values of the G rows go to one W x G LDS tile, then the whole block stores
// it (G contiguous y values per kx).  Two block-wide syncs per field instead
// of two per E/2 rounds.
// ----------------------------------------------------------------------------
template <typename T, typename Plan, int G, bool SPLIT> constexpr int tx_fwd_w_lds() {
  constexpr int stage = 2 * (Plan::N / 2 + 1) * tile_pitch<G, T>();
  constexpr int fft = (SPLIT ? 1 : 2) * t_plane<Plan, G>();
  return stage > fft ? stage : fft;
}

template <typename T, typename Plan, int G, bool SPLIT, bool NT = false>
__global__ void __launch_bounds__(G *Plan::TL) k_tx_fwd_w(TStepArgs<T> a) {
  __shared__ T lds[tx_fwd_w_lds<T, Plan, G, SPLIT>()];
  constexpr int TL = Plan::TL, E = Plan::E, N = Plan::N;
  const int t = threadIdx.x % TL, g = threadIdx.x / TL;
  const int ngy = (a.ny + G - 1) / G;
  const unsigned bid = xcd_block<T, G>(blockIdx.x, gridDim.x);
  const int z = (int)bid / ngy, y0 = ((int)bid % ngy) * G;
  const int y = y0 + g;
  const bool active = y < a.ny;
  const T *__restrict__ row = a.psi_in + ((long)z * a.ny + (active ? y : y0)) * N;

  cplx<T> v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const T u = ld_val<NT>(row + (t + TL * e));
    v[e] = cplx<T>{u, nl_value<Plan>(a.nl, u)};
  }
  fwd_split_store_wide<T, Plan, G, SPLIT, NT>(v, lds, t, g, z, y0, a);
}


// ----------------------------------------------------------------------------
// y passes with few-round transpositions: the N positions of the G lines go
// through the LDS in NR rounds of N/NR rows ([position][line] tile), one
// block-wide store (forward) or load (inverse) per round instead of one per
// value of E.
// ----------------------------------------------------------------------------
template <typename T, typename Plan, int G, bool SPLIT, int NR> constexpr int ty_w_lds() {
  constexpr int stage = 2 * (Plan::N / NR) * tile_pitch<G, T>();
  constexpr int fft = (SPLIT ? 1 : 2) * t_plane<Plan, G>();
  return stage > fft ? stage : fft;
}

template <typename T, typename Plan, int G, bool SPLIT, int NR, bool NT = false>
__global__ void __launch_bounds__(G *Plan::TL) k_ty_fwd_w(TStepArgs<T> a) {
  __shared__ T lds[ty_w_lds<T, Plan, G, SPLIT, NR>()];
  constexpr int TL = Plan::TL, E = Plan::E, N = Plan::N, RP = row_padded(N);
  constexpr int PL = t_plane<Plan, G>();
  constexpr int ROWS = N / NR, EPR = E / NR; // tile rows, values per thread per round
  static_assert(E % NR == 0, "NR must divide E");
  const int t = threadIdx.x % TL, g = threadIdx.x / TL;
  const int ngz = (a.nz + G - 1) / G;
  const unsigned bid = xcd_block<T, G>(blockIdx.x, gridDim.x);
  const int kx = (int)bid / ngz, z0 = ((int)bid % ngz) * G;
  const int z = z0 + g;
  const bool active = z < a.nz;
  const cplx<T> *__restrict__ in = blockIdx.y ? a.B1 : a.A1;
  cplx<T> *__restrict__ out = blockIdx.y ? a.B2 : a.A2;
  const cplx<T> *line = in + ((long)(active ? z : z0) * a.W + kx) * N;

  cplx<T> v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = ld_line<NT>(line + (t + TL * e));
  const LdsLine<T> ll{lds, g * RP, 1, PL};
  LineFft<T, Plan, -1, SPLIT, true, row_wave<Plan>()>::run(v, t, ll, a.tw_y);
  // ky = TL*e + t; round h holds e in [h*EPR, (h+1)*EPR), i.e. ky in [h*ROWS, (h+1)*ROWS)
  T *tr = lds, *ti = lds + ROWS * tile_pitch<G, T>();
#pragma unroll
  for (int h = 0; h < NR; ++h) {
    __syncthreads(); // previous users of the LDS are done
#pragma unroll
    for (int q = 0; q < EPR; ++q) {
      const int row = TL * q + t; // ky - h*ROWS
      tr[tile_idx<G, T>(row, g)] = v[h * EPR + q].x;
      ti[tile_idx<G, T>(row, g)] = v[h * EPR + q].y;
    }
    __syncthreads();
    if constexpr (sizeof(T) == 8) {
      tile_store<ROWS, G, G * TL, NT>(tr, ti, a.nz - z0, [&](int row, int gg) {
        return out + (((long)kx * N + h * ROWS + row) * a.nz + z0 + gg);
      });
    } else {
      for (int i = threadIdx.x; i < ROWS * G; i += G * TL) {
        const int row = i / G, gg = i % G;
        if (z0 + gg < a.nz)
          st_line<NT>(out + (((long)kx * N + h * ROWS + row) * a.nz + z0 + gg),
                      cplx<T>{tr[tile_idx<G, T>(row, gg)], ti[tile_idx<G, T>(row, gg)]});
      }
    }
  }
}



template <typename T, typename Plan, int G, bool SPLIT, int NR, bool NT = false>
__global__ void __launch_bounds__(G *Plan::TL) k_ty_inv_w(TStepArgs<T> a) {
  __shared__ T lds[ty_w_lds<T, Plan, G, SPLIT, NR>()];
  constexpr int TL = Plan::TL, E = Plan::E, N = Plan::N, RP = row_padded(N);
  constexpr int PL = t_plane<Plan, G>();
  constexpr int ROWS = N / NR, EPR = E / NR;
  static_assert(E % NR == 0, "NR must divide E");
  const int t = threadIdx.x % TL, g = threadIdx.x / TL;
  const int ngz = (a.nz + G - 1) / G;
  const unsigned bid = xcd_block<T, G>(blockIdx.x, gridDim.x);
  const int kx = (int)bid / ngz, z0 = ((int)bid % ngz) * G;
  const int z = z0 + g;

  cplx<T> v[E];
  T *tr = lds, *ti = lds + ROWS * tile_pitch<G, T>();
#pragma unroll
  for (int h = 0; h < NR; ++h) {
    const cplx<T> *__restrict__ src = a.A2 + ((long)kx * N + h * ROWS) * a.nz + z0;
    tile_load<ROWS, G, G * TL, NT>(tr, ti, a.nz - z0,
                                   [&](int row, int gg) { return src + ((long)row * a.nz + gg); });
    __syncthreads();
#pragma unroll
    for (int q = 0; q < EPR; ++q) {
      const int row = TL * q + t;
      v[h * EPR + q] = cplx<T>{tr[tile_idx<G, T>(row, g)], ti[tile_idx<G, T>(row, g)]};
    }
    __syncthreads(); // the tile is refilled (or becomes the FFT exchange area)
  }
  const LdsLine<T> ll{lds, g * RP, 1, PL};
  LineFft<T, Plan, +1, SPLIT, true, row_wave<Plan>()>::run(v, t, ll, a.tw_y);
  if (z < a.nz) {
    cplx<T> *__restrict__ line = a.A1 + ((long)z * a.W + kx) * N;
#pragma unroll
    for (int e = 0; e < E; ++e) st_line<NT>(line + (t + TL * e), v[e]);
  }
}


// ----------------------------------------------------------------------------
// Step boundary of a multi-step PFC run, fused: the inverse x pass of step n
// (T1 rows of Psi -> psi, nx/2-point plan PH) and the forward x pass of step
// n+1 (psi -> A1: Psi, B1: N, nx-point plan PF) for G consecutive rows of
// plane z, psi never leaving the block.  The C2R leaves psi as the pairs
// (psi[2m], psi[2m+1]) of m = t + TL*e; one LDS exchange hands each thread
// the positions t + TL*e of the nx-point ownership, then the forward x runs
// as in k_tx_fwd_w.  PH and PF use the same threads per line.  Reads C and
// writes 2C: 2R less than the two passes it replaces.  A block reads and then
// rewrites only its own (z, y-group) tile of A1.
// ----------------------------------------------------------------------------
template <typename T, typename PH, typename PF, int G, bool SPLIT> constexpr int tx_fuse_lds() {
  constexpr int inv = tx_inv_lds<T, PH, G, true>();
  constexpr int xch = G * row_padded(PF::N);
  constexpr int fwd = tx_fwd_w_lds<T, PF, G, SPLIT>();
  return inv > xch ? (inv > fwd ? inv : fwd) : (xch > fwd ? xch : fwd);
}

// WPE > 1 caps the registers so that WPE waves fit on a SIMD.
template <typename T, typename PH, typename PF, int G, bool SPLIT, int WPE = 1>
__global__ void __launch_bounds__(G *PF::TL, WPE) k_tx_fuse(TStepArgs<T> a) {
  static_assert(PH::TL == PF::TL && PF::N == 2 * PH::N, "paired plans");
  constexpr int TL = PH::TL, EH = PH::E, EF = PF::E, M = PH::N;
  constexpr int RPH = row_padded(M), RPF = row_padded(PF::N);
  constexpr int PLH = t_plane<PH, G>();
  constexpr int W = M + 1;
  __shared__ T lds[tx_fuse_lds<T, PH, PF, G, SPLIT>()];
  const int t = threadIdx.x % TL, g = threadIdx.x / TL;
  const int ngy = (a.ny + G - 1) / G;
  const unsigned bid = xcd_block<T, G>(blockIdx.x, gridDim.x);
  const int z = (int)bid / ngy, y0 = ((int)bid % ngy) * G;

  // inverse x: transposed load of the G rows, C2R merge, inverse transform
  cplx<T> v[EH];
  {
    T *sr = lds, *si = lds + W * tile_pitch<G, T>();
    const cplx<T> *__restrict__ src = a.A1 + (long)z * a.W * a.ny + y0;
    tile_load<W, G, G * TL, false>(sr, si, a.ny - y0,
                                   [&](int kx, int gg) { return src + ((long)kx * a.ny + gg); });
    __syncthreads();
#pragma unroll
    for (int e = 0; e < EH; ++e) {
      const int k = t + TL * e;
      v[e] = c2r_merge(cplx<T>{sr[tile_idx<G, T>(k, g)], si[tile_idx<G, T>(k, g)]},
                       cplx<T>{sr[tile_idx<G, T>(M - k, g)], si[tile_idx<G, T>(M - k, g)]},
                       a.wn[k], k == 0);
    }
    __syncthreads(); // the staging area becomes the FFT exchange area
  }
  const LdsLine<T> llh{lds, g * RPH, 1, PLH};
  LineFft<T, PH, +1, false, true, row_wave<PH>()>::run(v, t, llh, a.tw_xm);

  // psi to the nx-point ownership, then psi + i*N(psi)
  __syncthreads();
  T *row = lds + g * RPF;
#pragma unroll
  for (int e = 0; e < EH; ++e) {
    // positions 2m and 2m + 1 of m = t + TL*e: pad_pos(2m) = row_slot<2TL>(2t, e)
    // (2m mod 16 <= 14, so 2m + 1 lands in the next slot)
    const int i = row_slot<2 * TL>(2 * t, e);
    row[i] = v[e].x;
    row[i + 1] = v[e].y;
  }
  __syncthreads();
  cplx<T> w[EF];
#pragma unroll
  for (int e = 0; e < EF; ++e) {
    const T u = row[row_slot<TL>(t, e)];
    w[e] = cplx<T>{u, a.nl(u)};
  }
  __syncthreads(); // the LDS becomes the forward transform's exchange area
  fwd_split_store_wide<T, PF, G, SPLIT>(w, lds, t, g, z, y0, a);
}

} // namespace pfc_amd
