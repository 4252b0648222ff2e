/* CPU oracle for the BLS12-381 Fr prover operators -- TEST INFRASTRUCTURE.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this file's library; the product path never does.  See oracle.h for what
 * each function restates and how tests pin it.
 *
 * Arithmetic is 4 x 64-bit limbs with unsigned __int128 (CIOS Montgomery),
 * deliberately a different limb size and algorithm from the GPU's 8 x 32-bit
 * FIPS product so that the two cannot share a mistake.
 */
#define _GNU_SOURCE
#include "oracle.h"
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

typedef unsigned __int128 u128;

/* r, PLONK/utils/mont/cpu/ff/bls12-381.hpp:7-11 */
static const uint64_t P[4] = {0xffffffff00000001ull, 0x53bda402fffe5bfeull,
                              0x3339d80809a1d805ull, 0x73eda753299d7d48ull};
/* -r^-1 mod 2^64, cpu/ff/bls12-381.hpp:24 */
static const uint64_t N0 = 0xfffffffeffffffffull;
/* R^2 mod r, cpu/ff/bls12-381.hpp:12-16 */
static const uint64_t RR[4] = {0xc999e990f3f29c6dull, 0x2b6cedcb87925c23ull,
                               0x05d314967254398full, 0x0748d9d99f59ff11ull};
/* R mod r, cpu/ff/bls12-381.hpp:17-21 */
static const uint64_t ONE[4] = {0x00000001fffffffeull, 0x5884b7fa00034802ull,
                                0x998c4fefecbc4ff5ull, 0x1824b159acc5056full};

static int geq_p(const uint64_t* a) {
  for (int i = 3; i >= 0; i--) {
    if (a[i] > P[i]) return 1;
    if (a[i] < P[i]) return 0;
  }
  return 1;
}

static void sub_p(uint64_t* a) {
  uint64_t br = 0;
  for (int i = 0; i < 4; i++) {
    u128 d = (u128)a[i] - P[i] - br;
    a[i] = (uint64_t)d;
    br = (uint64_t)(d >> 64) & 1;
  }
}

int or_is_canonical(const uint64_t* a) { return !geq_p(a); }

void or_add(uint64_t* o, const uint64_t* a, const uint64_t* b) {
  uint64_t t[4];
  uint64_t c = 0;
  for (int i = 0; i < 4; i++) {
    u128 s = (u128)a[i] + b[i] + c;
    t[i] = (uint64_t)s;
    c = (uint64_t)(s >> 64);
  }
  if (c || geq_p(t)) sub_p(t);
  memcpy(o, t, 32);
}

void or_sub(uint64_t* o, const uint64_t* a, const uint64_t* b) {
  uint64_t t[4];
  uint64_t br = 0;
  for (int i = 0; i < 4; i++) {
    u128 d = (u128)a[i] - b[i] - br;
    t[i] = (uint64_t)d;
    br = (uint64_t)(d >> 64) & 1;
  }
  if (br) {
    uint64_t c = 0;
    for (int i = 0; i < 4; i++) {
      u128 s = (u128)t[i] + P[i] + c;
      t[i] = (uint64_t)s;
      c = (uint64_t)(s >> 64);
    }
  }
  memcpy(o, t, 32);
}

void or_neg(uint64_t* o, const uint64_t* a) {
  static const uint64_t Z[4] = {0, 0, 0, 0};
  or_sub(o, Z, a);
}

/* CIOS Montgomery product a*b*2^-256 mod r */
void or_mul(uint64_t* o, const uint64_t* a, const uint64_t* b) {
  uint64_t t[6] = {0, 0, 0, 0, 0, 0};
  for (int i = 0; i < 4; i++) {
    uint64_t C = 0;
    for (int j = 0; j < 4; j++) {
      u128 s = (u128)a[j] * b[i] + t[j] + C;
      t[j] = (uint64_t)s;
      C = (uint64_t)(s >> 64);
    }
    u128 s = (u128)t[4] + C;
    t[4] = (uint64_t)s;
    t[5] = (uint64_t)(s >> 64);
    uint64_t m = t[0] * N0;
    s = (u128)m * P[0] + t[0];
    C = (uint64_t)(s >> 64);
    for (int j = 1; j < 4; j++) {
      s = (u128)m * P[j] + t[j] + C;
      t[j - 1] = (uint64_t)s;
      C = (uint64_t)(s >> 64);
    }
    s = (u128)t[4] + C;
    t[3] = (uint64_t)s;
    t[4] = t[5] + (uint64_t)(s >> 64);
  }
  if (t[4] || geq_p(t)) sub_p(t);
  memcpy(o, t, 32);
}

void or_one(uint64_t* o) { memcpy(o, ONE, 32); }

void or_to_mont(uint64_t* o, const uint64_t* a) { or_mul(o, a, RR); }

void or_from_mont(uint64_t* o, const uint64_t* a) {
  static const uint64_t U1[4] = {1, 0, 0, 0};
  or_mul(o, a, U1);
}

void or_from_u64(uint64_t* o, uint64_t v) {
  uint64_t t[4] = {v, 0, 0, 0};
  or_to_mont(o, t);
}

void or_pow(uint64_t* o, const uint64_t* a, const uint64_t* e) {
  uint64_t acc[4], base[4];
  memcpy(acc, ONE, 32);
  memcpy(base, a, 32);
  for (int i = 0; i < 256; i++) {
    if ((e[i >> 6] >> (i & 63)) & 1) or_mul(acc, acc, base);
    or_mul(base, base, base);
  }
  memcpy(o, acc, 32);
}

void or_inv(uint64_t* o, const uint64_t* a) {
  /* Fermat: a^(r-2); 0^(r-2) = 0 gives the reference's x/0 = 0
   * (mont_t.cuh:1117-1126) */
  uint64_t e[4] = {P[0] - 2, P[1], P[2], P[3]};
  or_pow(o, a, e);
}

void or_vec_add(uint64_t* o, const uint64_t* a, const uint64_t* b, uint64_t n) {
  for (uint64_t i = 0; i < n; i++) or_add(o + 4 * i, a + 4 * i, b + 4 * i);
}
void or_vec_sub(uint64_t* o, const uint64_t* a, const uint64_t* b, uint64_t n) {
  for (uint64_t i = 0; i < n; i++) or_sub(o + 4 * i, a + 4 * i, b + 4 * i);
}
void or_vec_mul(uint64_t* o, const uint64_t* a, const uint64_t* b, uint64_t n) {
  for (uint64_t i = 0; i < n; i++) or_mul(o + 4 * i, a + 4 * i, b + 4 * i);
}
void or_vec_inv(uint64_t* o, const uint64_t* a, uint64_t n) {
  for (uint64_t i = 0; i < n; i++) or_inv(o + 4 * i, a + 4 * i);
}

/* exp_mod: o[i] = a[i]^e; e = 0 gives one, e = 1 a copy
 * (PLONK/utils/mont/cpu/mont_arithmetic.cpp:66-81) */
void or_vec_exp(uint64_t* o, const uint64_t* a, uint64_t n, uint32_t e) {
  const uint64_t ee[4] = {e, 0, 0, 0};
  for (uint64_t i = 0; i < n; i++) {
    if (e == 1)
      memcpy(o + 4 * i, a + 4 * i, 32);
    else
      or_pow(o + 4 * i, a + 4 * i, ee);
  }
}

/* gen_sequence: o[i] = x^i as a running product, o[0] = one, o[1] = x
 * (poly_eval_kernel, PLONK/utils/mont/cuda/mont_arithmetic.cu:104-114) */
void or_sequence(uint64_t* o, const uint64_t* x, uint64_t n) {
  if (n == 0) return;
  memcpy(o, ONE, 32);
  if (n > 1) memcpy(o + 4, x, 32);
  for (uint64_t i = 2; i < n; i++) or_mul(o + 4 * i, o + 4 * (i - 1), x);
}

void or_group_gen(uint64_t* o) { or_from_u64(o, 7); }

/* omega_{2^32} = 7^((r-1)/2^32); omega_{2^lg} = omega_{2^32}^(2^(32-lg)).
 * Same closed form as forward_roots_of_unity in parameters/bls12_381.h. */
void or_root_of_unity(uint64_t* o, int lg) {
  uint64_t g[4];
  or_group_gen(g);
  /* (r-1) >> 32 */
  uint64_t rm1[4] = {P[0] - 1, P[1], P[2], P[3]};
  uint64_t e[4];
  for (int i = 0; i < 4; i++) {
    e[i] = rm1[i] >> 32;
    if (i < 3) e[i] |= rm1[i + 1] << 32;
  }
  uint64_t w[4];
  or_pow(w, g, e);
  for (int i = lg; i < 32; i++) or_mul(w, w, w);
  memcpy(o, w, 32);
}

void or_domain_size_inverse(uint64_t* o, int lg) {
  uint64_t two_k[4];
  uint64_t t[4] = {0, 0, 0, 0};
  t[lg >> 6] = 1ull << (lg & 63);
  or_to_mont(two_k, t);
  or_inv(o, two_k);
}

uint64_t or_bitrev(uint64_t v, int bits) {
  uint64_t r = 0;
  for (int i = 0; i < bits; i++) {
    r = (r << 1) | (v & 1);
    v >>= 1;
  }
  return r;
}

static uint64_t splitmix(uint64_t* s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

void or_fill_uniform(uint64_t* out, uint64_t n, uint64_t seed) {
  uint64_t s = seed;
  for (uint64_t i = 0; i < n; i++) {
    uint64_t l[4];
    do {
      for (int j = 0; j < 4; j++) l[j] = splitmix(&s);
      l[3] &= 0x7fffffffffffffffull;
    } while (geq_p(l));
    memcpy(out + 4 * i, l, 32);
  }
}

uint64_t or_seed(int op, int lg, int aux) {
  return 0x504C4E4B00000000ull ^ ((uint64_t)op << 40) ^ ((uint64_t)lg << 8) ^ (uint64_t)aux;
}

/* ---- host threads ------------------------------------------------------
 * The transforms below split their loops over host threads (pthreads).  The
 * arithmetic is exact, so the result does not depend on the thread count. */
static int g_threads = 0;

void or_set_threads(int t) { g_threads = t < 1 ? 1 : (t > 64 ? 64 : t); }

int or_get_threads(void) {
  if (g_threads > 0) return g_threads;
  const char* e = getenv("OR_THREADS");
  int t = e ? atoi(e) : (int)sysconf(_SC_NPROCESSORS_ONLN);
  if (t < 1) t = 1;
  if (t > 16) t = 16;
  return t;
}

typedef void (*range_fn)(void* ctx, uint64_t lo, uint64_t hi);
struct par_job {
  range_fn fn;
  void* ctx;
  uint64_t lo, hi;
};

static void* par_entry(void* p) {
  struct par_job* j = (struct par_job*)p;
  if (j->lo < j->hi) j->fn(j->ctx, j->lo, j->hi);
  return NULL;
}

/* fn(ctx, lo, hi) over [0, n) in contiguous ranges, one per thread; ranges
 * are multiples of `grain` except the last.  `cost`: elements touched per
 * index, so that small loops stay on the calling thread. */
static void par_for_cost(uint64_t n, uint64_t grain, uint64_t cost, range_fn fn, void* ctx) {
  int T = or_get_threads();
  if (grain == 0) grain = 1;
  uint64_t units = (n + grain - 1) / grain;
  if ((uint64_t)T > units) T = (int)units;
  if (T <= 1 || n * cost < 8192) {
    if (n) fn(ctx, 0, n);
    return;
  }
  pthread_t th[64];
  struct par_job jobs[64];
  for (int t = 0; t < T; t++) {
    uint64_t lo = units * t / T * grain, hi = units * (t + 1) / T * grain;
    if (hi > n) hi = n;
    jobs[t].fn = fn;
    jobs[t].ctx = ctx;
    jobs[t].lo = lo;
    jobs[t].hi = hi;
  }
  int started[64] = {0};
  for (int t = 1; t < T; t++) started[t] = pthread_create(&th[t], NULL, par_entry, &jobs[t]) == 0;
  par_entry(&jobs[0]);
  for (int t = 1; t < T; t++) {
    if (started[t])
      pthread_join(th[t], NULL);
    else
      par_entry(&jobs[t]); /* could not start a thread: run the range here */
  }
}

static void par_for(uint64_t n, uint64_t grain, range_fn fn, void* ctx) { par_for_cost(n, grain, 1, fn, ctx); }

/* ---- NTT: bit-reverse, then radix-2 DIT (natural in, natural out) ------ */
struct bitrev_ctx {
  uint64_t* d;
  int lg;
};
static void bitrev_range(void* c, uint64_t lo, uint64_t hi) {
  struct bitrev_ctx* x = (struct bitrev_ctx*)c;
  for (uint64_t i = lo; i < hi; i++) {
    uint64_t j = or_bitrev(i, x->lg);
    if (j > i) {
      uint64_t t[4];
      memcpy(t, x->d + 4 * i, 32);
      memcpy(x->d + 4 * i, x->d + 4 * j, 32);
      memcpy(x->d + 4 * j, t, 32);
    }
  }
}

/* o[i] = scale * base^(i + start) for i in [lo, hi) (each range starts from
 * one exponentiation, then steps by products) */
struct powers_ctx {
  uint64_t* o;
  const uint64_t* base;
  const uint64_t* scale; /* NULL: 1 */
  uint64_t start;
  int multiply; /* 1: o[i] *= scale * base^(i+start) in place */
};
static void powers_range(void* c, uint64_t lo, uint64_t hi) {
  struct powers_ctx* x = (struct powers_ctx*)c;
  uint64_t p[4], e[4] = {lo + x->start, 0, 0, 0};
  or_pow(p, x->base, e);
  if (x->scale) or_mul(p, p, x->scale);
  for (uint64_t i = lo; i < hi; i++) {
    if (x->multiply)
      or_mul(x->o + 4 * i, x->o + 4 * i, p);
    else
      memcpy(x->o + 4 * i, p, 32);
    or_mul(p, p, x->base);
  }
}

struct stage_ctx {
  uint64_t* d;
  const uint64_t* tw;
  uint64_t half;
  int s;    /* stage (half block 2^s) for stage_range */
  int lgB;  /* block_range: stages 0..lgB-1 on blocks of 2^lgB elements */
};
static void butterfly(uint64_t* a, uint64_t* b, const uint64_t* w) {
  uint64_t t[4];
  or_mul(t, b, w);
  or_sub(b, a, t);
  or_add(a, a, t);
}
/* stages 0 .. lgB-1 on each block of 2^lgB elements in [lo, hi) (block units) */
static void block_range(void* c, uint64_t lo, uint64_t hi) {
  struct stage_ctx* x = (struct stage_ctx*)c;
  const uint64_t B = 1ull << x->lgB;
  for (uint64_t b = lo; b < hi; b++) {
    uint64_t* d = x->d + 4 * b * B;
    for (int s = 0; s < x->lgB; s++) {
      uint64_t h = 1ull << s, stride = x->half >> s;
      for (uint64_t blk = 0; blk < B; blk += 2 * h)
        for (uint64_t j = 0; j < h; j++) butterfly(d + 4 * (blk + j), d + 4 * (blk + j + h), x->tw + 4 * (j * stride));
    }
  }
}
/* butterflies [lo, hi) of stage s (butterfly i pairs blk + j and blk + j + h) */
static void stage_range(void* c, uint64_t lo, uint64_t hi) {
  struct stage_ctx* x = (struct stage_ctx*)c;
  const int s = x->s;
  const uint64_t h = 1ull << s, stride = x->half >> s;
  for (uint64_t i = lo; i < hi; i++) {
    uint64_t blk = (i >> s) << (s + 1), j = i & (h - 1);
    butterfly(x->d + 4 * (blk + j), x->d + 4 * (blk + j + h), x->tw + 4 * (j * stride));
  }
}

static void ntt_with_root(uint64_t* d, int lg, const uint64_t* w) {
  uint64_t n = 1ull << lg;
  if (lg == 0) return;
  struct bitrev_ctx bc = {d, lg};
  par_for(n, 1, bitrev_range, &bc);
  /* tw[j] = w^j, j < n/2 */
  uint64_t half = n >> 1;
  uint64_t* tw = (uint64_t*)malloc(32 * half);
  struct powers_ctx pc = {tw, w, NULL, 0, 0};
  par_for(half, 1, powers_range, &pc);
  /* the first stages block by block (2^12 elements = 128 KiB stay in cache),
   * the rest one stage at a time over all butterflies */
  int lgB = lg < 12 ? lg : 12;
  struct stage_ctx sc = {d, tw, half, 0, lgB};
  par_for_cost(n >> lgB, 1, 1ull << lgB, block_range, &sc);
  for (int s = lgB; s < lg; s++) {
    sc.s = s;
    par_for(half, 1, stage_range, &sc);
  }
  free(tw);
}

void or_ntt(uint64_t* data, int lg, int inverse) {
  uint64_t w[4];
  or_root_of_unity(w, lg);
  if (inverse) or_inv(w, w);
  ntt_with_root(data, lg, w);
  if (inverse) {
    uint64_t ninv[4], one[4];
    or_domain_size_inverse(ninv, lg);
    memcpy(one, ONE, 32);
    /* data[i] *= ninv * 1^i */
    struct powers_ctx pc = {data, one, ninv, 0, 1};
    par_for(1ull << lg, 1, powers_range, &pc);
  }
}

void or_coset_ntt(uint64_t* data, int lg, const uint64_t* shift) {
  /* data[i] *= shift^i */
  struct powers_ctx pc = {data, shift, NULL, 0, 1};
  par_for(1ull << lg, 1, powers_range, &pc);
  or_ntt(data, lg, 0);
}

struct strided_ctx {
  uint64_t* out;
  const uint64_t* in;
  uint64_t B, t, mask;
};
/* out[j*B + t] = in[j] */
static void interleave_range(void* c, uint64_t lo, uint64_t hi) {
  struct strided_ctx* x = (struct strided_ctx*)c;
  for (uint64_t j = lo; j < hi; j++) memcpy(x->out + 4 * (j * x->B + x->t), x->in + 4 * j, 32);
}
/* out[k] = in[k & mask] */
static void periodic_range(void* c, uint64_t lo, uint64_t hi) {
  struct strided_ctx* x = (struct strided_ctx*)c;
  for (uint64_t k = lo; k < hi; k++) memcpy(x->out + 4 * k, x->in + 4 * (k & x->mask), 32);
}

/* inverse coset NTT: x_i = shift^-i (1/N) sum_k X_k w_N^-ik, in place */
void or_coset_intt(uint64_t* data, int lg, const uint64_t* shift) {
  uint64_t inv[4];
  or_ntt(data, lg, 1);
  or_inv(inv, shift);
  struct powers_ctx pc = {data, inv, NULL, 0, 1};
  par_for(1ull << lg, 1, powers_range, &pc);
}

/* out[j*B + t] = coset_NTT_N(c; shift*w_M^t)[j], B = 2^lg_blowup */
void or_lde(const uint64_t* coeffs, uint64_t* out, int lg_n, int lg_blowup, const uint64_t* shift) {
  uint64_t n = 1ull << lg_n;
  uint64_t B = 1ull << lg_blowup;
  uint64_t wM[4];
  or_root_of_unity(wM, lg_n + lg_blowup);
  uint64_t* tmp = (uint64_t*)malloc(32 * n);
  uint64_t st[4];
  memcpy(st, shift, 32);
  for (uint64_t t = 0; t < B; t++) {
    memcpy(tmp, coeffs, 32 * n);
    or_coset_ntt(tmp, lg_n, st);
    struct strided_ctx ic = {out, tmp, B, t, 0};
    par_for(n, 1, interleave_range, &ic);
    or_mul(st, st, wM);
  }
  free(tmp);
}

/* w_3N = 7^((r-1)/(3*2^32)) squared (32-lg) times (3 divides r - 1) */
void or_root_of_unity_3n(uint64_t* o, int lg) {
  static const uint64_t RM1[4] = {0xffffffff00000000ull, 0x53bda402fffe5bfeull, 0x3339d80809a1d805ull,
                                  0x73eda753299d7d48ull};
  uint64_t e[4];
  for (int i = 0; i < 4; i++) e[i] = (RM1[i] >> 32) | (i < 3 ? RM1[i + 1] << 32 : 0);
  u128 rem = 0;
  for (int i = 3; i >= 0; i--) {
    u128 cur = (rem << 64) | e[i];
    e[i] = (uint64_t)(cur / 3);
    rem = cur % 3;
  }
  uint64_t seven[4];
  or_from_u64(seven, 7);
  or_pow(o, seven, e);
  for (int i = lg; i < 32; i++) or_mul(o, o, o);
}

/* blowup-3 LDE: out[3j + t] = coset_NTT_N(c; shift*w_3N^t)[j] */
void or_lde3(const uint64_t* coeffs, uint64_t* out, int lg_n, const uint64_t* shift) {
  uint64_t n = 1ull << lg_n;
  uint64_t w3[4];
  or_root_of_unity_3n(w3, lg_n);
  uint64_t* tmp = (uint64_t*)malloc(32 * n);
  uint64_t st[4];
  memcpy(st, shift, 32);
  for (uint64_t t = 0; t < 3; t++) {
    memcpy(tmp, coeffs, 32 * n);
    or_coset_ntt(tmp, lg_n, st);
    struct strided_ctx ic = {out, tmp, 3, t, 0};
    par_for(n, 1, interleave_range, &ic);
    or_mul(st, st, w3);
  }
  free(tmp);
}

void or_naive_lde(const uint64_t* coeffs, uint64_t* out, int lg_n, int lg_blowup) {
  uint64_t n = 1ull << lg_n;
  uint64_t M = n << lg_blowup;
  uint64_t* tmp = (uint64_t*)malloc(32 * n);
  memcpy(tmp, coeffs, 32 * n);
  or_ntt(tmp, lg_n, 0);
  struct strided_ctx pc = {out, tmp, 0, 0, n - 1};
  par_for(M, 1, periodic_range, &pc);
  free(tmp);
}

void or_grand_product(uint64_t* data, uint64_t n) {
  uint64_t z[4];
  memcpy(z, ONE, 32);
  for (uint64_t i = 0; i < n; i++) {
    uint64_t p[4];
    memcpy(p, data + 4 * i, 32);
    memcpy(data + 4 * i, z, 32);
    or_mul(z, z, p);
  }
}

void or_naive_gp(const uint64_t* in, uint64_t* out, uint64_t n) {
  if (n == 0) return;
  memset(out, 0, 32 * n);
  uint64_t st[4];
  memcpy(st, ONE, 32);
  memcpy(out, st, 32);
  for (uint64_t i = 0; i + 1 < n; i++) {
    or_mul(st, st, in + 4 * i);
    memcpy(out + 4 * i, st, 32);
  }
}

void or_eval(const uint64_t* coeffs, uint64_t n, const uint64_t* x, uint64_t* y) {
  uint64_t acc[4] = {0, 0, 0, 0};
  for (uint64_t i = n; i-- > 0;) {
    or_mul(acc, acc, x);
    or_add(acc, acc, coeffs + 4 * i);
  }
  memcpy(y, acc, 32);
}

void or_div(uint64_t* data, uint64_t n, const uint64_t* c, uint64_t* rem) {
  if (n == 0) return;
  uint64_t z[4];
  or_neg(z, c);
  uint64_t s[4];
  memcpy(s, data + 4 * (n - 1), 32);
  memset(data + 4 * (n - 1), 0, 32);
  for (uint64_t i = n - 1; i-- > 0;) {
    uint64_t a[4];
    memcpy(a, data + 4 * i, 32);
    memcpy(data + 4 * i, s, 32);
    or_mul(s, s, z);
    or_add(s, s, a);
  }
  if (rem) memcpy(rem, s, 32);
}

static int is_zero4(const uint64_t* a) { return (a[0] | a[1] | a[2] | a[3]) == 0; }

void or_naive_div(const uint64_t* in, uint64_t n, const uint64_t* c, const uint64_t* d1, uint64_t* q) {
  if (n < 2) return;
  uint64_t* rem = (uint64_t*)malloc(32 * n);
  memcpy(rem, in, 32 * n);
  memset(q, 0, 32 * (n - 1));
  uint64_t inv_lead[4];
  or_inv(inv_lead, d1);
  uint64_t len = n;
  while (len >= 2) {
    uint64_t cur[4], t[4];
    or_mul(cur, rem + 4 * (len - 1), inv_lead);
    memcpy(q + 4 * (len - 2), cur, 32);
    or_mul(t, cur, c);
    or_sub(rem + 4 * (len - 2), rem + 4 * (len - 2), t);
    or_mul(t, cur, d1);
    or_sub(rem + 4 * (len - 1), rem + 4 * (len - 1), t);
    /* the reference reads rem[-1] once every coefficient is zero
     * (miss_ops.cu:82-84); stop at an empty remainder instead */
    while (len > 0 && is_zero4(rem + 4 * (len - 1))) len--;
  }
  free(rem);
}

void or_perm_ratio(const uint64_t* a, const uint64_t* b, const uint64_t* c,
                   const uint64_t* s1, const uint64_t* s2, const uint64_t* s3,
                   const uint64_t* beta, const uint64_t* gamma,
                   const uint64_t* k1, const uint64_t* k2, const uint64_t* k3,
                   int lg_n, uint64_t* ratio) {
  uint64_t n = 1ull << lg_n;
  uint64_t w[4], x[4];
  or_root_of_unity(w, lg_n);
  memcpy(x, ONE, 32);
  const uint64_t* W[3] = {a, b, c};
  const uint64_t* S[3] = {s1, s2, s3};
  const uint64_t* K[3] = {k1, k2, k3};
  for (uint64_t i = 0; i < n; i++) {
    uint64_t num[4], den[4], t[4];
    memcpy(num, ONE, 32);
    memcpy(den, ONE, 32);
    for (int j = 0; j < 3; j++) {
      /* w_j + beta*k_j*omega^i + gamma */
      or_mul(t, beta, K[j]);
      or_mul(t, t, x);
      or_add(t, t, W[j] + 4 * i);
      or_add(t, t, gamma);
      or_mul(num, num, t);
      /* w_j + beta*s_j + gamma */
      or_mul(t, beta, S[j] + 4 * i);
      or_add(t, t, W[j] + 4 * i);
      or_add(t, t, gamma);
      or_mul(den, den, t);
    }
    or_inv(den, den);
    or_mul(ratio + 4 * i, num, den);
    or_mul(x, x, w);
  }
}

/* ---- segmented NTT step (stage_no_rotate) -------------------------------- */

/* bit_rev(i, nbits) of kernels.cuh:11-16 (32-bit index_t) */
static uint32_t rev32(uint32_t i, unsigned nbits) {
  uint32_t r = 0;
  for (int b = 0; b < 32; b++) r |= ((i >> b) & 1u) << (31 - b);
  return nbits ? r >> (32 - nbits) : 0;
}

/* o = omega_{2^lg}^e (roots cached; square-and-multiply over e's bits) */
static void root_pow(uint64_t* o, int lg, uint64_t e) {
  static uint64_t cache[33][4];
  static int have[33];
  if (!have[lg]) {
    or_root_of_unity(cache[lg], lg);
    have[lg] = 1;
  }
  uint64_t acc[4], b[4];
  or_one(acc);
  memcpy(b, cache[lg], 32);
  while (e) {
    if (e & 1) or_mul(acc, acc, b);
    e >>= 1;
    if (e) or_mul(b, b, b);
  }
  memcpy(o, acc, 32);
}

/* ntt.cu:168-178 */
void or_ct_stage_indices(uint64_t tid, int stage, int iterations, uint64_t* idx0, uint64_t* idx1) {
  const uint64_t inp_ntt_size = 1ull << stage;
  const uint64_t out_ntt_size = 1ull << (stage + iterations - 1);
  const uint64_t thread_ntt_pos = (tid & (out_ntt_size - 1)) >> (iterations - 1);
  uint64_t i0 = tid & ~(out_ntt_size - 1);
  i0 += (tid << stage) & (out_ntt_size - 1);
  i0 = i0 * 2 + thread_ntt_pos;
  *idx0 = i0;
  *idx1 = i0 + inp_ntt_size;
}

/* ntt.cu:190-195 */
uint64_t or_rotate_stage_output_index(uint64_t global_idx, int stage, int iterations) {
  const uint64_t mask = (1ull << (stage + iterations)) - (1ull << stage);
  uint64_t rotated = global_idx & mask;
  rotated = (rotated >> 1) | (rotated << (iterations - 1));
  return (global_idx & ~mask) | (rotated & mask);
}

int or_stage_no_rotate(uint64_t* chunk, uint64_t chunk_elems, int lg_domain, int iterations, int stage,
                       uint64_t chunk_id) {
  const int it = iterations;
  /* CTkernel_no_rotate: radix = max(6, it), cases 6, 7, 8, 10 */
  const int radix = it < 6 ? 6 : it;
  if (it < 1 || !(radix == 6 || radix == 7 || radix == 8 || radix == 10)) return -1;
  if (lg_domain > 28 || stage < 0 || stage + it > lg_domain) return -1;
  if (chunk_elems < 2 || (chunk_elems & (chunk_elems - 1)) || chunk_elems < (1ull << it)) return -1;
  const uint64_t nthreads = chunk_elems / 2;
  const uint64_t block = nthreads < (1ull << (radix - 1)) ? nthreads : (1ull << (radix - 1));
  /* offset = chunk_id * blockDim * gridDim, index_t arithmetic */
  const uint32_t offset = (uint32_t)(chunk_id * nthreads);
  uint64_t* r0 = (uint64_t*)malloc(nthreads * 32);
  uint64_t* r1 = (uint64_t*)malloc(nthreads * 32);
  uint64_t* x = (uint64_t*)malloc(nthreads * 32);
  /* intermediate_mul mode (ct_mixed_radix_wide.cuh:915-966): 0 at stage 0,
   * 2 (table of w_{2^L}^(row col), rows of 2^radix) at (6,6) L=12, (6,12)
   * L=18, (7,7) L=14, (8,8) L=16, else 1 (windowed w_{2^28} powers) */
  int mode = stage == 0 ? 0 : 1, L = 0;
  if (stage != 0) {
    if (radix == 6 && stage == 6) { mode = 2; L = 12; }
    if (radix == 6 && stage == 12) { mode = 2; L = 18; }
    if (radix == 7 && stage == 7) { mode = 2; L = 14; }
    if (radix == 8 && stage == 8) { mode = 2; L = 16; }
  }
  for (uint64_t t = 0; t < nthreads; t++) {
    const uint32_t gtid = (uint32_t)t + offset;
    memcpy(r0 + 4 * t, chunk + 4 * (2 * t), 32);
    memcpy(r1 + 4 * t, chunk + 4 * (2 * t + 1), 32);
    if (mode != 0) {
      const uint32_t diff_mask = (1u << (it - 1)) - 1;
      const uint32_t j = (gtid & diff_mask) * 2;
      const uint32_t gsize = 1u << (stage + it - 1);
      const uint32_t p = (gtid & (gsize - 1)) >> (it - 1);
      uint64_t f0[4], f1[4];
      if (mode == 1) {
        const unsigned nbits = 28 - stage;
        const uint32_t e0 = rev32(j, nbits) * p;
        const uint32_t e1 = p << (nbits - 1);
        uint64_t a[4];
        root_pow(f0, 28, e0);
        root_pow(a, 28, e1);
        or_mul(f1, a, f0); /* second_root *= first_root */
      } else {
        int sh = L - lg_domain;
        if (sh < 0) sh = 0;
        const unsigned nb = (unsigned)sh + (unsigned)it;
        const uint64_t i0 = ((uint64_t)p << radix) + rev32(j, nb);
        const uint64_t i1 = ((uint64_t)p << radix) + rev32(j + 1, nb);
        /* table[row * 2^radix + col] = w_{2^L}^(row col) */
        root_pow(f0, L, (i0 >> radix) * (i0 & ((1ull << radix) - 1)));
        root_pow(f1, L, (i1 >> radix) * (i1 & ((1ull << radix) - 1)));
      }
      or_mul(r0 + 4 * t, r0 + 4 * t, f0);
      or_mul(r1 + 4 * t, r1 + 4 * t, f1);
    }
    uint64_t tt[4];
    memcpy(tt, r1 + 4 * t, 32);
    or_sub(r1 + 4 * t, r0 + 4 * t, tt);
    or_add(r0 + 4 * t, r0 + 4 * t, tt);
  }
  uint64_t* tw = (uint64_t*)malloc((1ull << it) * 32);
  for (int s = 1; s < it; s++) {
    const uint64_t laneMask = 1ull << (s - 1), thrdMask = (1ull << s) - 1;
    /* d_radix6_twiddles[rank << (6 - (s+1))] (s < 6) or
     * radixX[rank << (radix - (s+1))]: w_{2^(s+1)}^rank either way */
    for (uint64_t k = 0; k <= thrdMask; k++) {
      if (s < 6)
        root_pow(tw + 4 * k, 6, k << (6 - (s + 1)));
      else
        root_pow(tw + 4 * k, radix, k << (radix - (s + 1)));
    }
    for (uint64_t t = 0; t < nthreads; t++) {
      const uint64_t rank = (t % block) & thrdMask;
      memcpy(x + 4 * t, rank < laneMask ? r1 + 4 * t : r0 + 4 * t, 32);
    }
    for (uint64_t t = 0; t < nthreads; t++) {
      const uint64_t rank = (t % block) & thrdMask;
      const int pos = rank < laneMask;
      const uint64_t* xp = x + 4 * (t ^ laneMask);
      if (pos)
        memcpy(r1 + 4 * t, xp, 32);
      else
        memcpy(r0 + 4 * t, xp, 32);
      uint64_t tt[4];
      or_mul(tt, tw + 4 * rank, r1 + 4 * t);
      or_sub(r1 + 4 * t, r0 + 4 * t, tt);
      or_add(r0 + 4 * t, r0 + 4 * t, tt);
    }
  }
  for (uint64_t t = 0; t < nthreads; t++) {
    memcpy(chunk + 4 * (2 * t), r0 + 4 * t, 32);
    memcpy(chunk + 4 * (2 * t + 1), r1 + 4 * t, 32);
  }
  free(tw);
  free(r0);
  free(r1);
  free(x);
  return 0;
}
