// NTT pass kernel template (included by the per-radix translation units).
// NTT / iNTT passes for gfx950.
//
// Algorithm: four-step (Bailey) decimation in frequency, applied recursively.
// A size-N transform, N = 2^lg, is split into passes of M' = 2^m points
// (3 <= m <= 6).  With N = N1*N2, j = j1 + N1*j2, k = k2 + N2*k1:
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
// Workgroup: 128 lanes own a tile of 1024 elements (8 per lane, all in
// VGPRs, about 110-128 VGPRs per lane, so four waves share each SIMD: one
// wave alone cannot issue multiply-adds back to back, and four reach ~15 %
// more than two).  A column of M' = 8*A points is computed as DFT_8 over
// rho2 (one per lane, in registers), an inner twiddle w_M'^(rho1 k2), one
// exchange through LDS (16 KiB per workgroup in two rounds), then DFT_A over
// rho1 (8/A per lane, in registers).  The tile holds C = 1024/M' adjacent
// columns, so every global access moves C*32 >= 512 contiguous bytes.  The
// in-register DFTs are radix-2 DIF networks whose twiddles (powers of w_16)
// are compile-time constants read from SGPRs.  Between passes the data is
// held as pbd words (fr29.hpp), not reduced; the last pass writes canonical
// values in [0, r).
#pragma once
#include "fr29.hpp"
#include "pl_internal.hpp"

namespace pl {

// w_16^e and w_16^-e, e = 0..7, as w R' mod r (R' = 2^261) in balanced
// digits (b_i in [-2^28, 2^28), sum b_i 2^(29 i)): the wave-uniform factor of
// the radix-2^29 product (fr29.hpp)
static __constant__ const uint32_t kW16b[2][8][9] = {
    {{0xffffffbau, 0x00000230u, 0xfcb61180u, 0x0a4e5c01u, 0x0ee8b1a2u, 0xf6e6aedfu, 0xf907f8bcu, 0x0853ddf8u, 0x004d043fu},
     {0x04056c96u, 0x0f64e4d8u, 0xfc1c7b56u, 0xfcbda468u, 0xf4e761dbu, 0xf28ae057u, 0xfd2f3145u, 0x0f463b6eu, 0x006dc1feu},
     {0x085dfee9u, 0x0f4995e4u, 0xf941c136u, 0x0b284958u, 0xfc313523u, 0xff28f6a6u, 0x07ad921eu, 0x0d72d344u, 0x002ca644u},
     {0xf81392abu, 0xf40c77eeu, 0x0175e03eu, 0x09f5a4e3u, 0x0f090202u, 0x0c6252cbu, 0xfc504ed8u, 0x038e921cu, 0x00408014u},
     {0xf139f60bu, 0xf0567553u, 0xfe6cfcdeu, 0x03679a02u, 0x08c51075u, 0xff9df144u, 0xfeb67e0cu, 0xfa6828a5u, 0x00231dc7u},
     {0x0142510bu, 0x02a002b3u, 0x09a9466eu, 0xff776142u, 0x002a3050u, 0xf0591c89u, 0x04f2cdf4u, 0x0d6e4f9eu, 0x006990bbu},
     {0xf6551b88u, 0xff9842ffu, 0x0f9a0689u, 0xfb208bceu, 0xfe704152u, 0x09d87034u, 0xfd460dfbu, 0x063990acu, 0x0050578fu},
     {0xf7ddaa4du, 0xfd2e4775u, 0x06b5ad2fu, 0xf9075af7u, 0x06ef6bbeu, 0xf25fb8efu, 0x07ff9caeu, 0xf5571281u, 0x00687c14u}},
    {{0xffffffbau, 0x00000230u, 0xfcb61180u, 0x0a4e5c01u, 0x0ee8b1a2u, 0xf6e6aedfu, 0xf907f8bcu, 0x0853ddf8u, 0x004d043fu},
     {0x082255b4u, 0x02d1b883u, 0xf8e15291u, 0x0240ab09u, 0xf690e97eu, 0xf9a44be2u, 0x0d21303au, 0xf50e212eu, 0x000b7194u},
     {0x09aae479u, 0x0067bcf9u, 0x0ffcf937u, 0x00277a31u, 0xff1013eau, 0x022b949du, 0xf7dabeecu, 0x042ba304u, 0x00239618u},
     {0xfebdaef6u, 0xfd5ffd45u, 0xf5edb952u, 0xfbd0a4beu, 0xfd5624ecu, 0xfbaae848u, 0xf02dfef4u, 0xfcf6e412u, 0x000a5cecu},
     {0x0ec609f6u, 0x0fa98aa5u, 0x012a02e2u, 0xf7e06bfeu, 0xf4bb44c7u, 0x0c66138du, 0xf66a4edbu, 0x0ffd0b0bu, 0x0050cfe0u},
     {0x07ec6d56u, 0x0bf3880au, 0xfe211f82u, 0xf152611du, 0x0e77533au, 0xffa1b205u, 0xf8d07e0fu, 0x06d6a194u, 0x00336d93u},
     {0xf7a20118u, 0xf0b66a14u, 0x06553e8au, 0xf01fbca8u, 0x014f2019u, 0x0cdb0e2bu, 0x0d733ac9u, 0xfcf2606bu, 0x00474763u},
     {0xfbfa936bu, 0xf09b1b20u, 0x037a846au, 0xfe8a6198u, 0x0898f361u, 0xf979247au, 0xf7f19ba3u, 0xfb1ef842u, 0x00062ba9u}}};

template <bool INV>
__device__ __forceinline__ void w16b(int e, uint32_t (&w)[9]) {
#pragma unroll
  for (int i = 0; i < 9; i++) w[i] = kW16b[INV ? 1 : 0][e][i];
}

// an ordering point for the eight limbs of v (no instruction)
__device__ __forceinline__ void pin(fr& v) {
  asm volatile("" : "+v"(v.l[0]), "+v"(v.l[1]), "+v"(v.l[2]), "+v"(v.l[3]), "+v"(v.l[4]), "+v"(v.l[5]),
               "+v"(v.l[6]), "+v"(v.l[7]));
}

__host__ __device__ constexpr int brev(int v, int bits) {
  int r = 0;
  for (int i = 0; i < bits; i++) r |= ((v >> i) & 1) << (bits - 1 - i);
  return r;
}

// Radix-2 DIF DFT of size 2^LOGB on x[OFF + i], natural order in,
// bit-reversed order out: afterwards x[OFF + brev(k)] = X[k].  Elements are
// signed-limb values (v9): sums and differences are limb-wise, and with every
// input limb a balanced digit (|l| <= 2^28) each output limb is a signed sum
// of at most 2^LOGB balanced digits or product digits (below 2^29), which
// stays within 32 bits for LOGB <= 3 (the first term of every such sum enters
// with a plus sign, so the extremes are 2^31 - 1 and -2^31).
template <int LOGB, bool INV, int OFF, int NX>
__device__ __forceinline__ void dft_dif(v9 (&x)[NX]) {
  static_assert(LOGB <= 3, "limb growth: at most three butterfly levels");
  constexpr int B = 1 << LOGB;
  static_for<0, LOGB>([&](auto st_) {
    constexpr int st = decltype(st_)::value;
    constexpr int h = B >> (st + 1);
    static_for<0, (1 << st)>([&](auto blk_) {
      constexpr int s = decltype(blk_)::value * 2 * h;
      static_for<0, h>([&](auto i_) {
        constexpr int i = decltype(i_)::value;
        // one butterfly at a time: the empty asm statements order the
        // butterflies, so the live set stays at the lane's elements plus one
        // butterfly's temporaries
        v9 a = x[OFF + s + i];
        v9 b = x[OFF + s + i + h];
        pin9(a);
        pin9(b);
        v9 sum = v9_add(a, b);
        v9 d = v9_sub(a, b);
        if constexpr (i != 0) {
          uint32_t w[9];
          w16b<INV>((i * 8) / h, w);
          d = v9_mul<true, false, false>(d, w);
        }
        pin9(sum);
        pin9(d);
        x[OFF + s + i] = sum;
        x[OFF + s + i + h] = d;
      });
    });
  });
}

struct PassArgs {
  const fr* in;
  fr* out;
  const fr* premul;     // !LAST: optional factor per input element [(poly & pm_mask) << lgN | pos], pbd words of c R''
  const fr* tw;         // !LAST: pass twiddles (pbd words of w R' mod r), index j + S*k within the sub-problem
  const fr* inner;      // w_256^e (or w_256^-e), e < 256, stored form (small-size kernel)
  const fr* inner29;    // the same as pbd words of w R' mod r (pass kernels)
  uint64_t in_pstride;  // element stride between input polys (0: shared input)
  uint64_t nlines;      // lines (columns) in this launch
  uint32_t pm_mask;
  uint32_t lgN;         // log2 poly size
  uint32_t lgMp;        // log2 sub-problem size of this pass
  uint32_t lgS;         // log2 column stride (0 in the last pass)
  uint32_t lgI;         // LAST: output interleave
  uint32_t ndig;        // LAST: number of earlier passes
  uint32_t do_scale;    // multiply every output by `scale` (single-pass transforms)
  uint32_t in_ext;      // the input is caller data (any 256-bit values), not a previous pass's pbd words
  uint32_t lgrad[PL_MAX_PASSES];  // LAST: log2 radices of the earlier passes, pass 1 first
  fr scale;             // stored form (small-size kernel)
  fr scale29;           // pbd word of scale R' mod r (pass kernels)
};

// Tile geometry (compile-time; see the file header):
//   PL_NTT_LGPER   log2 elements per lane (3: 8 elements, DFT_8 in registers)
//   PL_NTT_LGTILE  log2 elements per workgroup tile (128 lanes x 8)
//   PL_NTT_PHASES  LDS exchange rounds (9/PHASES limbs each)
//   PL_NTT_WAVES   waves per SIMD the register allocation must allow
// Passes have 2^m-point columns, m = PL_NTT_LGPER + LA, LA <= min(4, LGPER).
#ifndef PL_NTT_LGTILE
#define PL_NTT_LGTILE 10
#endif
#ifndef PL_NTT_LGPER
#define PL_NTT_LGPER 3
#endif
constexpr int kLgPerT = PL_NTT_LGPER;
constexpr int kPerT = 1 << kLgPerT;
constexpr int kPmG = kPerT / 4;  // factor-table loads in flight per group (first pass of an LDE)
#ifndef PL_NTT_PHASES
#define PL_NTT_PHASES 3
#endif
#ifndef PL_NTT_WAVES
#define PL_NTT_WAVES 4
#endif
constexpr int kLgTile = PL_NTT_LGTILE;
constexpr int kTile = 1 << kLgTile;
constexpr int kThreads = kTile / kPerT;
constexpr int kPhases = PL_NTT_PHASES;
constexpr int kW = 9 / kPhases;  // limbs per exchange round
static_assert(kW * kPhases == 9, "the exchange rounds cover the nine limbs");

template <int W> struct lds_vec;
template <> struct lds_vec<4> { typedef uint4 type; };
template <> struct lds_vec<3> { typedef uint3 type; };
template <> struct lds_vec<2> { typedef uint2 type; };
template <> struct lds_vec<1> { typedef uint32_t type; };

template <int W>
__device__ __forceinline__ typename lds_vec<W>::type pack_limbs(const uint32_t* v) {
  if constexpr (W == 4) return make_uint4(v[0], v[1], v[2], v[3]);
  else if constexpr (W == 3) return make_uint3(v[0], v[1], v[2]);
  else if constexpr (W == 2) return make_uint2(v[0], v[1]);
  else return v[0];
}
template <int W>
__device__ __forceinline__ void unpack_limbs(uint32_t* v, const typename lds_vec<W>::type& q) {
  if constexpr (W == 4) { v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w; }
  else if constexpr (W == 3) { v[0] = q.x; v[1] = q.y; v[2] = q.z; }
  else if constexpr (W == 2) { v[0] = q.x; v[1] = q.y; }
  else v[0] = q;
}

// One pass over columns of M' = 8*A points, A = 2^LA; blk = tile index.
//
// Arithmetic: the lane's elements are signed-limb values (v9).  The input is
// unpacked to balanced digits (caller data: any 256-bit value; later passes:
// the previous pass's pbd words), so each in-register DFT starts from
// balanced digits and its limbs stay within 32 bits; every product takes
// such limbs.  Value bounds (V = input magnitude, at most 2^256 < 2.21 r): a
// DFT_8 output is below 8 V = 17.7 r; an inner-twiddle product of it lies in
// (-1.25 r, 0.25 r) (x w / 2^261 < 0.0142 x for w < r); element k2 = 0, which
// has no inner twiddle, is reduced into (-0.51 r, 0.51 r); so a DFT_A output
// is below 10 r, and the pass twiddle product with centred digits
// (mont29 BM) lies in (-0.65 r, 0.65 r): its balanced digits fit a pbd word
// (top digit below 2^23).  The last pass reduces its outputs to [0, r).
// SCALE: multiply every output by a.scale29 (single-pass inverse transforms;
// a separate instance, so the other last-pass kernels keep no registers for it)
template <int LA, bool INV, bool LAST, bool SCALE>
__device__ __forceinline__ void ntt_pass_body(const PassArgs& a, const uint64_t blk) {
  static_assert(!SCALE || LAST, "the scale factor is applied by the last pass");
  constexpr int A = 1 << LA;
  constexpr int m = LA + kLgPerT;
  constexpr int LGC = kLgTile - m;  // log2 lines per tile
  constexpr int C = 1 << LGC;
  constexpr int F2 = kPerT / A;   // DFT_A groups per thread in sub-step 2
  typedef typename lds_vec<kW>::type vec;
  __shared__ vec lds[kTile];  // kW limbs of every tile element

  const int t = threadIdx.x;
  const uint64_t line0 = blk << LGC;
  const uint32_t lgL = a.lgN - m;  // log2 lines per polynomial = log2(N / M')

  // ---- sub-step 1: kPerT points of one (line, rho1), DFT over rho2 ----------
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
  v9 x[kPerT];
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
    // balanced digits: caller data is any 256-bit value (optionally times a
    // premultiplication factor), a previous pass's output is a pbd word.  Each
    // branch loads and unpacks on its own, so no loaded word lives across it.
    if (a.in_ext) {
      if (!LAST && a.premul) {
        const fr* pm = a.premul + ((uint64_t)(poly & a.pm_mask) << a.lgN) + pos0 + (uint64_t)rho1 * estride;
        const fr* sp = src + pos0 + (uint64_t)rho1 * estride;
        // factors are pbd words of c R' (the coset powers); the product with
        // balanced output digits lies in (-r, 0.03 r) and leaves balanced
        // digits, as the DFT wants.  kPmG elements and factors at a time: each
        // group's loads start after the previous group's products (the asm
        // statements order them), so few loaded words are held beside the
        // lane's results; addresses are opaque bases plus wave-uniform offsets
        static_for<0, kPerT / kPmG>([&](auto g_) {
          constexpr int g = decltype(g_)::value;
          const fr* q = pm;
          const fr* d = sp;
          asm volatile("" : "+v"(q), "+v"(d));
          fr f[kPmG], v[kPmG];
          static_for<0, kPmG>([&](auto r_) {
            constexpr int r = decltype(r_)::value;
            v[r] = fr_load(d + (uint64_t)(A * (kPmG * g + r)) * estride);
            f[r] = fr_load(q + (uint64_t)(A * (kPmG * g + r)) * estride);
          });
          static_for<0, kPmG>([&](auto r_) {
            constexpr int r = decltype(r_)::value;
            constexpr int k = kPmG * g + r;
            v9 xx;
            unpack_ext(v[r].l, xx.l);
            uint32_t w[9];
            pbd_unpack(f[r].l, w);
            x[k] = v9_mul<false, false, true>(xx, w);
            pin9(x[k]);
          });
        });
      } else {
        fr raw[kPerT];
        static_for<0, kPerT>([&](auto r2_) {
          constexpr int r2 = decltype(r2_)::value;
          raw[r2] = fr_load(src + pos0 + (uint64_t)(rho1 + A * r2) * estride);
        });
        static_for<0, kPerT>([&](auto r2_) {
          constexpr int r2 = decltype(r2_)::value;
          unpack_ext(raw[r2].l, x[r2].l);
        });
      }
    } else {
      static_for<0, kPerT>([&](auto r2_) {
        constexpr int r2 = decltype(r2_)::value;
        const fr raw = fr_load(src + pos0 + (uint64_t)(rho1 + A * r2) * estride);
        pbd_unpack(raw.l, x[r2].l);
      });
    }
  } else {
    static_for<0, kPerT>([&](auto r2_) {
#pragma unroll
      for (int i = 0; i < 9; i++) x[decltype(r2_)::value].l[i] = 0;
    });
  }
  dft_dif<kLgPerT, INV, 0>(x);  // x[brev3(k2)] = sum_rho2 z[rho1 + A rho2] w8^(rho2 k2)
  if constexpr (LA > 0) {
    // k2 = 0 takes no twiddle: back to balanced digits of a small value, so
    // that DFT_A again starts from balanced digits (the products' outputs
    // below are balanced, mont29 BT)
    {
      v9 t0;
      reduce_centered(x[0].l, t0.l);
      x[0] = t0;
    }
    static_for<1, kPerT>([&](auto k2_) {
      constexpr int k2 = decltype(k2_)::value;
      const int e = rho1 * k2;  // < M'
      pin9(x[brev(k2, kLgPerT)]);
      uint32_t w[9];
      pbd_unpack(fr_load(a.inner29 + (e << (8 - m))).l, w);
      x[brev(k2, kLgPerT)] = v9_mul<false, false, true>(x[brev(k2, kLgPerT)], w);
    });
  }
  // Exchange through LDS in kPhases rounds of kW limbs (3 rounds of 12 bytes
  // = 12 KiB for a 1024-element tile, so eight workgroups = four waves per
  // SIMD fit a CU).  LDS slot of W[line][rho1][k2]; consecutive lanes hit
  // consecutive slots.  Each round's gathered limbs go back into the
  // registers that round vacated: y lives in x (x[f*A + r1] afterwards).
  // The line index is XOR-swizzled with rho1 so that both the line-major
  // (intermediate passes) and the rho-major (last pass) write mappings and
  // the line-major reads are free of LDS bank conflicts.
  // a fresh copy of the thread index, so that sub-step 2's indices are not
  // shared with sub-step 1's and kept live across it
  int t2 = t;
  asm volatile("" : "+v"(t2));
  const int lc2 = t2 & (C - 1);
  const int g = t2 >> LGC;
  v9 (&y)[kPerT] = x;
  static_for<0, kPhases>([&](auto h_) {
    constexpr int h = decltype(h_)::value;
    uint32_t hv[kPerT][kW];
    static_for<0, kPerT>([&](auto k2_) {
      constexpr int k2 = decltype(k2_)::value;
#pragma unroll
      for (int w = 0; w < kW; w++) hv[k2][w] = x[brev(k2, kLgPerT)].l[kW * h + w];
    });
    if constexpr (h > 0) __syncthreads();  // previous round's reads done before overwrite
    static_for<0, kPerT>([&](auto k2_) {
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
  // no barrier follows: lanes of lines beyond the launch leave here
  const uint64_t line2 = line0 + lc2;
  if (line2 >= a.nlines) return;
  // last pass: the lane's output base, computed before DFT_A so that its
  // 64-bit temporaries are not live beside the eight results
  fr* base = nullptr;
  const uint32_t lgI = a.lgI;
  if constexpr (LAST) {
    uint64_t poly, lp;
    if (lgI) {
      poly = line2 & ((1ull << lgI) - 1);
      lp = line2 >> lgI;
    } else {
      poly = line2 >> lgL;
      lp = line2 & ((1ull << lgL) - 1);
    }
    fr* dst = a.out + ((poly >> lgI) << (a.lgN + lgI)) + (poly & ((1ull << lgI) - 1));
    base = dst + (lp << lgI);
    asm volatile("" : "+v"(base));
  }
  if constexpr (LA > 0) {
    static_for<0, F2>([&](auto f_) { dft_dif<LA, INV, decltype(f_)::value * A>(y); });
  }
  // y[f*A + brev_LA(k1)] = Z[k2 + kPerT*k1], k2 = g*F2 + f

  if (LAST) {
    if constexpr (SCALE) {
      // single-pass inverse transforms: the factor 2^-lg, once per element
      uint32_t w[9];
      pbd_unpack(a.scale29.l, w);
      static_for<0, kPerT>([&](auto i_) {
        constexpr int i = decltype(i_)::value;
        pin9(y[i]);
        y[i] = v9_mul<false, false, false>(y[i], w);
      });
    }
    static_for<0, F2>([&](auto f_) {
      constexpr int f = decltype(f_)::value;
      const int k2 = g * F2 + f;
      static_for<0, A>([&](auto k1_) {
        constexpr int k1 = decltype(k1_)::value;
        const uint64_t k = (uint64_t)(k2 + kPerT * k1);
        v9 v = y[f * A + brev(k1, LA)];
        pin9(v);
        uint32_t tt[9];
        reduce_below_2r(v.l, tt);
        fr o;
        join29(tt, o.l);
        // the address from an opaque copy of the lane's base, so that the
        // eight addresses are not all computed ahead and held
        fr* d = base;
        asm volatile("" : "+v"(d));
        fr c = fr_reduce_once(o);
        pin(c);  // one element's reduction at a time
        fr_store(d + (k << (lgL + lgI)), c);
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
        const uint64_t k = (uint64_t)(k2 + kPerT * k1);
        v9 v = y[f * A + brev(k1, LA)];
        pin9(v);
        uint32_t w[9];
        pbd_unpack(fr_load(tw + (k << a.lgS)).l, w);
        v = v9_mul<false, true, true>(v, w);
        fr o;
        pbd_pack(v.l, o.l);
        fr_store(dst + (k << a.lgS), o);
      });
    });
  }
}


template <int LA, bool INV, bool LAST, bool SCALE>
__global__ void __launch_bounds__(kThreads, PL_NTT_WAVES) ntt_pass_kernel(PassArgs a) {
  ntt_pass_body<LA, INV, LAST, SCALE>(a, blockIdx.x);
}

// Ragged form: one launch runs the tiles of many polynomials of different
// sizes (one PassArgs per segment).  seg_block0[s] is segment s's first
// block, ascending; a block finds its segment by binary search (uniform
// scalar loads, a few steps for the tens of segments of a batch).
template <int LA, bool INV, bool LAST, bool SCALE>
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
  ntt_pass_body<LA, INV, LAST, SCALE>(a, b - seg_block0[lo]);
}

// Host launcher for one radix: LA = log2(column size) - kLgPerT.
template <int LA, bool INV, bool LAST, bool SCALE = false>
static void launch_pass(const PassArgs& a, hipStream_t s) {
  constexpr int m = LA + kLgPerT;
  uint64_t tiles = (a.nlines + (kTile >> m) - 1) >> (kLgTile - m);
  hipLaunchKernelGGL((ntt_pass_kernel<LA, INV, LAST, SCALE>), dim3((unsigned)tiles), dim3(kThreads), 0, s, a);
}

template <int LA, bool INV, bool LAST, bool SCALE = false>
static void launch_pass_ragged(const PassArgs* segs, const uint32_t* seg_block0, uint32_t nsegs, uint32_t nblocks,
                               hipStream_t s) {
  hipLaunchKernelGGL((ntt_pass_ragged_kernel<LA, INV, LAST, SCALE>), dim3(nblocks), dim3(kThreads), 0, s, segs,
                     seg_block0, nsegs);
}

// scale (a.do_scale) only with inv && last
#define PL_DEFINE_PASS_LAUNCHER(LA)                                                            \
  void launch_pass_la##LA(bool inv, bool last, const PassArgs& a, hipStream_t s) {             \
    if (inv) {                                                                                 \
      if (last) {                                                                              \
        if (a.do_scale) launch_pass<LA, true, true, true>(a, s);                               \
        else launch_pass<LA, true, true>(a, s);                                                \
      } else {                                                                                 \
        launch_pass<LA, true, false>(a, s);                                                    \
      }                                                                                        \
    } else {                                                                                   \
      if (last) launch_pass<LA, false, true>(a, s);                                            \
      else launch_pass<LA, false, false>(a, s);                                                \
    }                                                                                          \
  }                                                                                            \
  void launch_pass_ragged_la##LA(bool inv, bool last, bool scale, const PassArgs* segs, const uint32_t* sb, \
                                 uint32_t ns, uint32_t nb, hipStream_t s) {                    \
    if (inv) {                                                                                 \
      if (last) {                                                                              \
        if (scale) launch_pass_ragged<LA, true, true, true>(segs, sb, ns, nb, s);              \
        else launch_pass_ragged<LA, true, true>(segs, sb, ns, nb, s);                          \
      } else {                                                                                 \
        launch_pass_ragged<LA, true, false>(segs, sb, ns, nb, s);                              \
      }                                                                                        \
    } else {                                                                                   \
      if (last) launch_pass_ragged<LA, false, true>(segs, sb, ns, nb, s);                      \
      else launch_pass_ragged<LA, false, false>(segs, sb, ns, nb, s);                          \
    }                                                                                          \
  }

#define PL_DECLARE_RAGGED(LA)                                                                  \
  void launch_pass_ragged_la##LA(bool inv, bool last, bool scale, const PassArgs* segs, const uint32_t* sb, \
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
