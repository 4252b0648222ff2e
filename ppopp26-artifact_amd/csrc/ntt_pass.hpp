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
// are compile-time constants read from SGPRs.
#pragma once
#include "fr29.hpp"
#include "pl_internal.hpp"

namespace pl {

// w_16^e and w_16^-e, e = 0..7, as w R' mod r (R' = 2^261) in nine 29-bit
// limbs: the constant operand of the radix-2^29 product (fr29.hpp)
static __constant__ const uint32_t kW16r[2][8][9] = {
    {{0x1fffffbau, 0x0000022fu, 0x1cb61180u, 0x0a4e5c00u, 0x0ee8b1a2u, 0x16e6aedfu, 0x1907f8bbu, 0x0853ddf7u, 0x004d043fu},
     {0x04056c96u, 0x0f64e4d8u, 0x1c1c7b56u, 0x1cbda467u, 0x14e761dau, 0x128ae056u, 0x1d2f3144u, 0x0f463b6du, 0x006dc1feu},
     {0x085dfee9u, 0x0f4995e4u, 0x1941c136u, 0x0b284957u, 0x1c313523u, 0x1f28f6a5u, 0x07ad921du, 0x0d72d344u, 0x002ca644u},
     {0x181392abu, 0x140c77edu, 0x0175e03du, 0x09f5a4e3u, 0x0f090202u, 0x0c6252cbu, 0x1c504ed8u, 0x038e921bu, 0x00408014u},
     {0x1139f60bu, 0x10567552u, 0x1e6cfcddu, 0x03679a01u, 0x08c51075u, 0x1f9df144u, 0x1eb67e0bu, 0x1a6828a4u, 0x00231dc6u},
     {0x0142510bu, 0x02a002b3u, 0x09a9466eu, 0x1f776142u, 0x002a304fu, 0x10591c89u, 0x04f2cdf3u, 0x0d6e4f9eu, 0x006990bbu},
     {0x16551b88u, 0x1f9842feu, 0x0f9a0688u, 0x1b208bceu, 0x1e704151u, 0x09d87033u, 0x1d460dfbu, 0x063990abu, 0x0050578fu},
     {0x17ddaa4du, 0x1d2e4774u, 0x06b5ad2eu, 0x19075af7u, 0x06ef6bbdu, 0x125fb8efu, 0x07ff9cadu, 0x15571281u, 0x00687c13u}},
    {{0x1fffffbau, 0x0000022fu, 0x1cb61180u, 0x0a4e5c00u, 0x0ee8b1a2u, 0x16e6aedfu, 0x1907f8bbu, 0x0853ddf7u, 0x004d043fu},
     {0x082255b4u, 0x02d1b883u, 0x18e15291u, 0x0240ab08u, 0x1690e97eu, 0x19a44be1u, 0x0d213039u, 0x150e212eu, 0x000b7193u},
     {0x09aae479u, 0x0067bcf9u, 0x0ffcf937u, 0x00277a31u, 0x1f1013eau, 0x022b949cu, 0x17dabeecu, 0x042ba303u, 0x00239618u},
     {0x1ebdaef6u, 0x1d5ffd44u, 0x15edb951u, 0x1bd0a4bdu, 0x1d5624ebu, 0x1baae847u, 0x102dfef3u, 0x1cf6e411u, 0x000a5cebu},
     {0x0ec609f6u, 0x0fa98aa5u, 0x012a02e2u, 0x17e06bfeu, 0x14bb44c6u, 0x0c66138cu, 0x166a4edbu, 0x0ffd0b0au, 0x0050cfe0u},
     {0x07ec6d56u, 0x0bf3880au, 0x1e211f82u, 0x1152611cu, 0x0e775339u, 0x1fa1b205u, 0x18d07e0eu, 0x06d6a193u, 0x00336d93u},
     {0x17a20118u, 0x10b66a13u, 0x06553e89u, 0x101fbca8u, 0x014f2018u, 0x0cdb0e2bu, 0x0d733ac9u, 0x1cf2606bu, 0x00474762u},
     {0x1bfa936bu, 0x109b1b1fu, 0x037a8469u, 0x1e8a6198u, 0x0898f360u, 0x1979247au, 0x17f19ba2u, 0x1b1ef841u, 0x00062ba8u}}};

template <bool INV>
__device__ __forceinline__ void w16r(int e, uint32_t (&w)[9]) {
#pragma unroll
  for (int i = 0; i < 9; i++) w[i] = kW16r[INV ? 1 : 0][e][i];
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
// bit-reversed order out: afterwards x[OFF + brev(k)] = X[k].  Values are kept
// lazily in [0, 2r) (fr_add_lazy / fr_sub_lazy; the radix-2^29 product returns
// below 1.03 r); the pass reduces them once before storing.
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
        // one butterfly at a time: the empty asm statements order the
        // butterflies, so the live set stays at the lane's elements plus one
        // butterfly's temporaries
        fr a = x[OFF + s + i];
        fr b = x[OFF + s + i + h];
        pin(a);
        pin(b);
        fr sum = fr_add_lazy(a, b);
        fr d = fr_sub_lazy(a, b);
        if constexpr (i != 0) {
          uint32_t w[9];
          w16r<INV>((i * 8) / h, w);
          d = fr_mul29<true>(d, w);
        }
        pin(sum);
        pin(d);
        x[OFF + s + i] = sum;
        x[OFF + s + i + h] = d;
      });
    });
  });
}

struct PassArgs {
  const fr* in;
  fr* out;
  const fr* premul;     // !LAST: optional factor per input element [(poly & pm_mask) << lgN | pos]
  const fr* tw;         // !LAST: pass twiddles (w R' mod r), index j + S*k within the sub-problem
  const fr* inner;      // w_256^e (or w_256^-e), e < 256, stored form (small-size kernel)
  const fr* inner29;    // the same as w R' mod r (pass kernels)
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
  fr scale;             // stored form (small-size kernel)
  fr scale29;           // scale R' mod r (pass kernels)
};

// Tile geometry (compile-time; see the file header):
//   PL_NTT_LGPER   log2 elements per lane (3: 8 elements, DFT_8 in registers)
//   PL_NTT_LGTILE  log2 elements per workgroup tile (128 lanes x 8)
//   PL_NTT_PHASES  LDS exchange rounds (8/PHASES limbs each)
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
#define PL_NTT_PHASES 2
#endif
#ifndef PL_NTT_WAVES
#define PL_NTT_WAVES 4
#endif
constexpr int kLgTile = PL_NTT_LGTILE;
constexpr int kTile = 1 << kLgTile;
constexpr int kThreads = kTile / kPerT;
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
  fr x[kPerT];
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
    static_for<0, kPerT>([&](auto r2_) {
      constexpr int r2 = decltype(r2_)::value;
      x[r2] = fr_load(src + pos0 + (uint64_t)(rho1 + A * r2) * estride);
    });
    if (!LAST && a.premul) {
      const fr* pm = a.premul + ((uint64_t)(poly & a.pm_mask) << a.lgN) + pos0;
      // kPmG factors at a time: each group's loads start after the previous
      // group's products (the asm statements order them), so at most kPmG
      // factors are held in registers beside the lane's elements
      static_for<0, kPerT / kPmG>([&](auto g_) {
        constexpr int g = decltype(g_)::value;
        const fr* q = pm;
        asm volatile("" : "+v"(q));
        fr f[kPmG];
        static_for<0, kPmG>([&](auto r_) {
          constexpr int r = decltype(r_)::value;
          f[r] = fr_load(q + (uint64_t)(rho1 + A * (kPmG * g + r)) * estride);
        });
        static_for<0, kPmG>([&](auto r_) {
          constexpr int r = decltype(r_)::value;
          x[kPmG * g + r] = fr_mul(x[kPmG * g + r], f[r]);
          pin(x[kPmG * g + r]);
        });
      });
    }
  } else {
    static_for<0, kPerT>([&](auto r2_) { x[decltype(r2_)::value] = fr_zero(); });
  }
  dft_dif<kLgPerT, INV, 0>(x);  // x[brev4(k2)] = sum_rho2 z[rho1 + A rho2] w16^(rho2 k2)
  if constexpr (LA > 0) {
    static_for<1, kPerT>([&](auto k2_) {
      constexpr int k2 = decltype(k2_)::value;
      const int e = rho1 * k2;  // < M'
      pin(x[brev(k2, kLgPerT)]);
      x[brev(k2, kLgPerT)] = fr_mul29(x[brev(k2, kLgPerT)], fr_load(a.inner29 + (e << (8 - m))));
    });
  }
  // Exchange through LDS in kPhases rounds of kW limbs (2 rounds of 16 bytes
  // = 16 KiB for a 1024-element tile, so eight workgroups = four waves per
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
  fr (&y)[kPerT] = x;
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
  if constexpr (LA > 0) {
    static_for<0, F2>([&](auto f_) { dft_dif<LA, INV, decltype(f_)::value * A>(y); });
  }
  // y[f*A + brev_LA(k1)] = Z[k2 + kPerT*k1], k2 = g*F2 + f

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
        const uint64_t k = (uint64_t)(k2 + kPerT * k1);
        fr v = y[f * A + brev(k1, LA)];
        if (a.do_scale) v = fr_mul29(v, a.scale29);
        fr_store(dst + ((lp + (k << lgL)) << lgI), fr_reduce_once(v));
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
        fr v = y[f * A + brev(k1, LA)];
        pin(v);
        v = fr_mul29(v, fr_load(tw + (k << a.lgS)));
        if (a.do_scale) v = fr_mul29(v, a.scale29);
        fr_store(dst + (k << a.lgS), fr_reduce_once(v));
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
  constexpr int m = LA + kLgPerT;
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
