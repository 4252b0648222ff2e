// NTT pass kernel template (included by the per-radix translation units).
// NTT / iNTT passes for gfx950.
//
// Algorithm: four-step (Bailey) decimation in frequency, applied recursively.
// A size-N transform, N = 2^lg, is split into passes of M' = 2^m points
// (4 <= m <= 8).  With N = N1*N2, j = j1 + N1*j2, k = k2 + N2*k1:
//
//   X[k2 + N2 k1] = sum_j1 w_N1^(j1 k1) * ( w_N^(j1 k2) * sum_j2 x[j1 + N1 j2] w_N2^(j2 k2) )
//
// Pass 1 does the inner column DFTs (size N2 = 2^m, stride N1) and the twiddle
// w_N^(j1 k2), in place; the rows of length N1 that remain are contiguous and
// are split the same way by the next pass.  The last pass transforms
// contiguous sub-rows and writes each result straight to its digit-reversed
// (natural-order) position, so no separate permutation pass exists.  This
// replaces the reference's host bit-reversal plus sppark-style CT passes
// (ntt.cu:979-1037; ntt_kernel/algorithm.cuh:8-59; ct_mixed_radix_wide.cuh:47-189)
// with the same mathematical result.
//
// Workgroup: 128 threads own a tile of 2048 elements (16 per thread, all in
// VGPRs).  A column of M' = 16*A points is computed as DFT_16 over rho2 (one
// per thread, in registers), an inner twiddle w_M'^(rho1 k2), one exchange
// through LDS, then DFT_A over rho1 (16/A per thread, in registers).  The
// tile holds C = 2048/M' adjacent columns, so every global access moves
// C*32 >= 256 contiguous bytes.  The in-register DFTs are radix-2 DIF
// networks whose twiddles (powers of w_16) are compile-time constants read
// from SGPRs.
#pragma once
#include "pl_internal.hpp"

namespace pl {

// w_16^e and w_16^-e, e = 0..7, Montgomery form
static __constant__ const uint32_t kW16[2][8][8] = {
    {{0xfffffffeu, 0x00000001u, 0x00034802u, 0x5884b7fau, 0xecbc4ff5u, 0x998c4fefu, 0xacc5056fu, 0x1824b159u},
     {0x88202b65u, 0x1acf64e4u, 0x819f004fu, 0x5b18b637u, 0x3107fdf5u, 0xee649cabu, 0xce9012a0u, 0x27a8543du},
     {0x2c42eff8u, 0x6ecf4995u, 0x4d5dfa57u, 0x6ec1fa83u, 0x5656b15cu, 0x3ff5cf38u, 0x229e031bu, 0x54b80267u},
     {0x45c09c96u, 0x5fb40c77u, 0x8b8b1b1bu, 0xfff59b78u, 0x6b7eed47u, 0x023ba68bu, 0x76f3e130u, 0x4e17f66fu},
     {0xaa89cfb1u, 0xf3b05674u, 0x6006b9feu, 0x072f0140u, 0x25667a26u, 0xce9a0dbfu, 0x2d598374u, 0x4d2ce405u},
     {0x0b0a1289u, 0x25c2a002u, 0x7d082188u, 0x971c717fu, 0xd81ac5d8u, 0x4739dccau, 0xeee341c1u, 0x4f607ba9u},
     {0x3eb2a8ddu, 0x111f9842u, 0x6f38b840u, 0x1f772784u, 0x8e3ca13cu, 0xb28758a3u, 0xed9a48f0u, 0x5974f9f7u},
     {0xdcbeed53u, 0x0ddd2e46u, 0x33dbdd55u, 0xdf977d7fu, 0xc1470bbau, 0xdda3b36au, 0xb67a22b2u, 0x4818fbf6u}},
    {{0xfffffffeu, 0x00000001u, 0x00034802u, 0x5884b7fau, 0xecbc4ff5u, 0x998c4fefu, 0xacc5056fu, 0x1824b159u},
     {0x234112aeu, 0xf222d1b8u, 0xcc227ea9u, 0x74262683u, 0x485acc4au, 0x5596249du, 0x73235a95u, 0x2bd4ab5cu},
     {0xc14d5724u, 0xeee067bcu, 0x90c5a3beu, 0x34467c7eu, 0x7b6536c9u, 0x80b27f64u, 0x3c033457u, 0x1a78ad5bu},
     {0xf4f5ed78u, 0xda3d5ffcu, 0x82f63a76u, 0xbca13283u, 0x3187122cu, 0xebfffb3du, 0x3aba3b86u, 0x248d2ba9u},
     {0x55763050u, 0x0c4fa98au, 0x9ff7a200u, 0x4c8ea2c2u, 0xe43b5ddfu, 0x649fca48u, 0xfc43f9d3u, 0x26c0c34du},
     {0xba3f636bu, 0xa04bf387u, 0x747340e3u, 0x53c8088au, 0x9e22eabdu, 0x30fe317cu, 0xb2a99c18u, 0x25d5b0e3u},
     {0xd3bd1009u, 0x9130b669u, 0xb2a061a7u, 0xe4fba97fu, 0xb34b26a8u, 0xf34408cfu, 0x06ff7a2cu, 0x1f35a4ecu},
     {0x77dfd49cu, 0xe5309b1au, 0x7e5f5bafu, 0xf8a4edcbu, 0xd899da0fu, 0x44d53b5cu, 0x5b0d6aa7u, 0x4c455315u}}};

template <bool INV>
__device__ __forceinline__ fr w16c(int e) {
  fr w;
#pragma unroll
  for (int i = 0; i < 8; i++) w.l[i] = kW16[INV ? 1 : 0][e][i];
  return w;
}

__host__ __device__ constexpr int brev(int v, int bits) {
  int r = 0;
  for (int i = 0; i < bits; i++) r |= ((v >> i) & 1) << (bits - 1 - i);
  return r;
}

// Radix-2 DIF DFT of size 2^LOGB on x[OFF + i], natural order in,
// bit-reversed order out: afterwards x[OFF + brev(k)] = X[k].
template <int LOGB, bool INV, int OFF, int NX>
__device__ __forceinline__ void dft_dif(fr (&x)[NX]) {
  constexpr int B = 1 << LOGB;
  static_for<0, LOGB>([&](auto st_) {
    constexpr int st = decltype(st_)::value;
    constexpr int h = B >> (st + 1);
    static_for<0, (1 << st)>([&](auto blk_) {
      constexpr int s = decltype(blk_)::value * 2 * h;
      static_for<0, h>([&](auto i_) {
        constexpr int i = decltype(i_)::value;
        fr a = x[OFF + s + i];
        fr b = x[OFF + s + i + h];
        x[OFF + s + i] = fr_add(a, b);
        fr d = fr_sub(a, b);
        if constexpr (i != 0) d = fr_mul_u(d, w16c<INV>((i * 8) / h));
        x[OFF + s + i + h] = d;
      });
    });
  });
}

struct PassArgs {
  const fr* in;
  fr* out;
  const fr* premul;     // !LAST: optional factor per input element [(poly & pm_mask) << lgN | pos]
  const fr* tw;         // !LAST: pass twiddles, index j + S*k within the sub-problem
  const fr* inner;      // w_256^e (or w_256^-e), e < 256
  uint64_t in_pstride;  // element stride between input polys (0: shared input)
  uint64_t nlines;      // lines (columns) in this launch
  uint32_t pm_mask;
  uint32_t lgN;         // log2 poly size
  uint32_t lgMp;        // log2 sub-problem size of this pass
  uint32_t lgS;         // log2 column stride (0 in the last pass)
  uint32_t lgI;         // LAST: output interleave
  uint32_t ndig;        // LAST: number of earlier passes
  uint32_t do_scale;    // multiply every output by `scale`
  uint32_t lgrad[PL_MAX_PASSES];  // LAST: log2 radices of the earlier passes, pass 1 first
  fr scale;
};

// Tile geometry (compile-time; see the file header):
//   PL_NTT_LGTILE  log2 elements per workgroup tile (16 per thread)
//   PL_NTT_PHASES  LDS exchange rounds (8/PHASES limbs each)
//   PL_NTT_WAVES   waves per SIMD the register allocation must allow
#ifndef PL_NTT_LGTILE
#define PL_NTT_LGTILE 11
#endif
#ifndef PL_NTT_PHASES
#define PL_NTT_PHASES 2
#endif
#ifndef PL_NTT_WAVES
#define PL_NTT_WAVES 2
#endif
constexpr int kLgTile = PL_NTT_LGTILE;
constexpr int kTile = 1 << kLgTile;
constexpr int kThreads = kTile / 16;
constexpr int kPhases = PL_NTT_PHASES;
constexpr int kW = 8 / kPhases;  // limbs per exchange round

template <int W> struct lds_vec;
template <> struct lds_vec<4> { typedef uint4 type; };
template <> struct lds_vec<2> { typedef uint2 type; };
template <> struct lds_vec<1> { typedef uint32_t type; };

template <int W>
__device__ __forceinline__ typename lds_vec<W>::type pack_limbs(const uint32_t* v) {
  if constexpr (W == 4) return make_uint4(v[0], v[1], v[2], v[3]);
  else if constexpr (W == 2) return make_uint2(v[0], v[1]);
  else return v[0];
}
template <int W>
__device__ __forceinline__ void unpack_limbs(uint32_t* v, const typename lds_vec<W>::type& q) {
  if constexpr (W == 4) { v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w; }
  else if constexpr (W == 2) { v[0] = q.x; v[1] = q.y; }
  else v[0] = q;
}

// One pass over columns of M' = 16*A points, A = 2^LA; blk = tile index.
template <int LA, bool INV, bool LAST>
__device__ __forceinline__ void ntt_pass_body(const PassArgs& a, const uint64_t blk) {
  constexpr int A = 1 << LA;
  constexpr int m = LA + 4;
  constexpr int LGC = kLgTile - m;  // log2 lines per tile
  constexpr int C = 1 << LGC;
  constexpr int F2 = 16 / A;   // DFT_A groups per thread in sub-step 2
  typedef typename lds_vec<kW>::type vec;
  __shared__ vec lds[kTile];  // kW limbs of every tile element

  const int t = threadIdx.x;
  const uint64_t line0 = blk << LGC;
  const uint32_t lgL = a.lgN - m;  // log2 lines per polynomial = log2(N / M')

  // ---- sub-step 1: 16 points of one (line, rho1), DFT_16 over rho2 --------
  // rho = rho1 + A*rho2.  Intermediate passes map lanes to adjacent columns
  // (coalesced strided rows); the last pass reads contiguous sub-rows, so
  // lanes take adjacent rho1.
  int lc, rho1;
  if (LAST) {
    rho1 = t & (A - 1);
    lc = t >> LA;
  } else {
    lc = t & (C - 1);
    rho1 = t >> LGC;
  }
  const uint64_t line = line0 + lc;
  fr x[16];
  if (line < a.nlines) {
    uint64_t poly, pos0, estride;
    if (LAST) {
      uint64_t lp;
      if (a.lgI) {
        poly = line & ((1ull << a.lgI) - 1);
        lp = line >> a.lgI;
      } else {
        poly = line >> lgL;
        lp = line & ((1ull << lgL) - 1);
      }
      // the sub-problem whose outputs land at lp + (N/M')*k: undo the digit
      // reversal lp = d1 + M1*(d2 + M2*(...)) -> r = (((d1)*M2 + d2)*M3 + ...)
      uint64_t r = 0, q = lp;
      for (uint32_t p = 0; p < a.ndig; p++) {
        uint32_t lr = a.lgrad[p];
        r = (r << lr) | (q & ((1ull << lr) - 1));
        q >>= lr;
      }
      pos0 = r << m;
      estride = 1;
    } else {
      poly = line >> lgL;
      uint64_t lp = line & ((1ull << lgL) - 1);
      uint64_t j = lp & ((1ull << a.lgS) - 1);
      uint64_t rr = lp >> a.lgS;
      pos0 = (rr << a.lgMp) + j;
      estride = 1ull << a.lgS;
    }
    const fr* src = a.in + poly * a.in_pstride;
    static_for<0, 16>([&](auto r2_) {
      constexpr int r2 = decltype(r2_)::value;
      x[r2] = fr_load(src + pos0 + (uint64_t)(rho1 + A * r2) * estride);
    });
    if (!LAST && a.premul) {
      const fr* pm = a.premul + ((uint64_t)(poly & a.pm_mask) << a.lgN);
      static_for<0, 16>([&](auto r2_) {
        constexpr int r2 = decltype(r2_)::value;
        x[r2] = fr_mul(x[r2], fr_load(pm + pos0 + (uint64_t)(rho1 + A * r2) * estride));
      });
    }
  } else {
    static_for<0, 16>([&](auto r2_) { x[decltype(r2_)::value] = fr_zero(); });
  }
  dft_dif<4, INV, 0>(x);  // x[brev4(k2)] = sum_rho2 z[rho1 + A rho2] w16^(rho2 k2)
  if constexpr (LA > 0) {
    static_for<1, 16>([&](auto k2_) {
      constexpr int k2 = decltype(k2_)::value;
      const int e = rho1 * k2;  // < M'
      x[brev(k2, 4)] = fr_mul(x[brev(k2, 4)], fr_load(a.inner + (e << (8 - m))));
    });
  }
  // Exchange through LDS in kPhases rounds of kW limbs (2 rounds of 16 bytes
  // = 32 KiB for a 2048-element tile, so four workgroups = two waves per SIMD
  // fit a CU).  LDS slot of W[line][rho1][k2]; consecutive lanes hit
  // consecutive slots.  Each round's gathered limbs go back into the
  // registers that round vacated: y lives in x (x[f*A + r1] afterwards).
  // The line index is XOR-swizzled with rho1 so that both the line-major
  // (intermediate passes) and the rho-major (last pass) write mappings and
  // the line-major reads are free of LDS bank conflicts.
  const int lc2 = t & (C - 1);
  const int g = t >> LGC;
  fr (&y)[16] = x;
  static_for<0, kPhases>([&](auto h_) {
    constexpr int h = decltype(h_)::value;
    uint32_t hv[16][kW];
    static_for<0, 16>([&](auto k2_) {
      constexpr int k2 = decltype(k2_)::value;
#pragma unroll
      for (int w = 0; w < kW; w++) hv[k2][w] = x[brev(k2, 4)].l[kW * h + w];
    });
    if constexpr (h > 0) __syncthreads();  // previous round's reads done before overwrite
    static_for<0, 16>([&](auto k2_) {
      constexpr int k2 = decltype(k2_)::value;
      const int slot = ((k2 * A + rho1) << LGC) + (lc ^ (rho1 & (C - 1)));
      lds[slot] = pack_limbs<kW>(hv[k2]);
    });
    __syncthreads();
    // sub-step 2 gathers: DFT_A over rho1 for F2 values of k2 of one line
    static_for<0, F2>([&](auto f_) {
      constexpr int f = decltype(f_)::value;
      const int k2 = g * F2 + f;
      static_for<0, A>([&](auto r1_) {
        constexpr int r1 = decltype(r1_)::value;
        const int slot = ((k2 * A + r1) << LGC) + (lc2 ^ (r1 & (C - 1)));
        unpack_limbs<kW>(&y[f * A + r1].l[kW * h], lds[slot]);
      });
    });
  });
  if constexpr (LA > 0) {
    static_for<0, F2>([&](auto f_) { dft_dif<LA, INV, decltype(f_)::value * A>(y); });
  }
  // y[f*A + brev_LA(k1)] = Z[k2 + 16*k1], k2 = g*F2 + f

  const uint64_t line2 = line0 + lc2;
  if (line2 >= a.nlines) return;
  if (LAST) {
    uint64_t poly, lp;
    if (a.lgI) {
      poly = line2 & ((1ull << a.lgI) - 1);
      lp = line2 >> a.lgI;
    } else {
      poly = line2 >> lgL;
      lp = line2 & ((1ull << lgL) - 1);
    }
    const uint32_t lgI = a.lgI;
    fr* dst = a.out + ((poly >> lgI) << (a.lgN + lgI)) + (poly & ((1ull << lgI) - 1));
    static_for<0, F2>([&](auto f_) {
      constexpr int f = decltype(f_)::value;
      const int k2 = g * F2 + f;
      static_for<0, A>([&](auto k1_) {
        constexpr int k1 = decltype(k1_)::value;
        const uint64_t k = (uint64_t)(k2 + 16 * k1);
        fr v = y[f * A + brev(k1, LA)];
        if (a.do_scale) v = fr_mul_u(v, a.scale);
        fr_store(dst + ((lp + (k << lgL)) << lgI), v);
      });
    });
  } else {
    const uint64_t poly = line2 >> lgL;
    const uint64_t lp = line2 & ((1ull << lgL) - 1);
    const uint64_t j = lp & ((1ull << a.lgS) - 1);
    const uint64_t rr = lp >> a.lgS;
    fr* dst = a.out + (poly << a.lgN) + (rr << a.lgMp) + j;
    const fr* tw = a.tw + j;
    static_for<0, F2>([&](auto f_) {
      constexpr int f = decltype(f_)::value;
      const int k2 = g * F2 + f;
      static_for<0, A>([&](auto k1_) {
        constexpr int k1 = decltype(k1_)::value;
        const uint64_t k = (uint64_t)(k2 + 16 * k1);
        fr v = y[f * A + brev(k1, LA)];
        v = fr_mul(v, fr_load(tw + (k << a.lgS)));
        if (a.do_scale) v = fr_mul_u(v, a.scale);
        fr_store(dst + (k << a.lgS), v);
      });
    });
  }
}


template <int LA, bool INV, bool LAST>
__global__ void __launch_bounds__(kThreads, PL_NTT_WAVES) ntt_pass_kernel(PassArgs a) {
  ntt_pass_body<LA, INV, LAST>(a, blockIdx.x);
}

// Ragged form: one launch runs the tiles of many polynomials of different
// sizes (one PassArgs per segment).  seg_block0[s] is segment s's first
// block, ascending; a block finds its segment by binary search (uniform
// scalar loads, a few steps for the tens of segments of a batch).
template <int LA, bool INV, bool LAST>
__global__ void __launch_bounds__(kThreads, PL_NTT_WAVES)
    ntt_pass_ragged_kernel(const PassArgs* __restrict__ segs, const uint32_t* __restrict__ seg_block0,
                           uint32_t nsegs) {
  const uint32_t b = blockIdx.x;
  uint32_t lo = 0, hi = nsegs - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (seg_block0[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  // read the arguments in place; a private copy would put the dynamically
  // indexed lgrad[] in scratch
  const PassArgs& a = segs[lo];
  ntt_pass_body<LA, INV, LAST>(a, b - seg_block0[lo]);
}

// Host launcher for one radix: LA = log2(column size) - 4.
template <int LA, bool INV, bool LAST>
static void launch_pass(const PassArgs& a, hipStream_t s) {
  constexpr int m = LA + 4;
  uint64_t tiles = (a.nlines + (kTile >> m) - 1) >> (kLgTile - m);
  hipLaunchKernelGGL((ntt_pass_kernel<LA, INV, LAST>), dim3((unsigned)tiles), dim3(kThreads), 0, s, a);
}

template <int LA, bool INV, bool LAST>
static void launch_pass_ragged(const PassArgs* segs, const uint32_t* seg_block0, uint32_t nsegs, uint32_t nblocks,
                               hipStream_t s) {
  hipLaunchKernelGGL((ntt_pass_ragged_kernel<LA, INV, LAST>), dim3(nblocks), dim3(kThreads), 0, s, segs, seg_block0,
                     nsegs);
}

#define PL_DEFINE_PASS_LAUNCHER(LA)                                        \
  void launch_pass_la##LA(bool inv, bool last, const PassArgs& a, hipStream_t s) { \
    if (inv) {                                                             \
      if (last) launch_pass<LA, true, true>(a, s);                         \
      else launch_pass<LA, true, false>(a, s);                             \
    } else {                                                               \
      if (last) launch_pass<LA, false, true>(a, s);                        \
      else launch_pass<LA, false, false>(a, s);                            \
    }                                                                      \
  }                                                                        \
  void launch_pass_ragged_la##LA(bool inv, bool last, const PassArgs* segs, const uint32_t* sb, \
                                 uint32_t ns, uint32_t nb, hipStream_t s) {                   \
    if (inv) {                                                             \
      if (last) launch_pass_ragged<LA, true, true>(segs, sb, ns, nb, s);   \
      else launch_pass_ragged<LA, true, false>(segs, sb, ns, nb, s);       \
    } else {                                                               \
      if (last) launch_pass_ragged<LA, false, true>(segs, sb, ns, nb, s);  \
      else launch_pass_ragged<LA, false, false>(segs, sb, ns, nb, s);      \
    }                                                                      \
  }

#define PL_DECLARE_RAGGED(LA)                                                                  \
  void launch_pass_ragged_la##LA(bool inv, bool last, const PassArgs* segs, const uint32_t* sb, \
                                 uint32_t ns, uint32_t nb, hipStream_t s);
// Kernel arguments of pass p of the call c when every pass is a regular
// tile pass (false for the tiny and single-pass premultiplied forms, which
// ntt_launch handles with extra kernels).
bool pass_args_regular(const Domain& d, const NttCall& c, int p, PassArgs& a);
PL_DECLARE_RAGGED(0)
PL_DECLARE_RAGGED(1)
PL_DECLARE_RAGGED(2)
PL_DECLARE_RAGGED(3)
PL_DECLARE_RAGGED(4)
void launch_pass_la0(bool inv, bool last, const PassArgs& a, hipStream_t s);
void launch_pass_la1(bool inv, bool last, const PassArgs& a, hipStream_t s);
void launch_pass_la2(bool inv, bool last, const PassArgs& a, hipStream_t s);
void launch_pass_la3(bool inv, bool last, const PassArgs& a, hipStream_t s);
void launch_pass_la4(bool inv, bool last, const PassArgs& a, hipStream_t s);

}  // namespace pl
