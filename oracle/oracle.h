/* CPU oracle for the BLS12-381 Fr prover operators.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline).  The product path never links or calls it.
 *
 * Every element is 4 little-endian uint64 limbs holding a canonical
 * Montgomery residue x*2^256 mod r, byte-identical to the reference's
 * cpu::BLS12_381_Fr_G1 (PLONK/utils/mont/cpu/ff/bls12-381.hpp:6-32) and
 * cuda::BLS12_381_Fr_G1 (PLONK/utils/mont/cuda/ff/bls12-381.hpp:7-45).
 *
 * How each function is pinned (tests/test_oracle.py):
 *   field ops       Python big-int closed form, 0/1/r-1 edge cases, and the
 *                   reference's own blst build in oracle/_ref when present.
 *   roots, 2^-k     the reference's constant tables (bls12_381.h), copied as
 *                   data into tests/golden/ref_constants.json.
 *   ntt / intt      brute-force O(N^2) DFT from Python big ints (lg <= 8),
 *                   iNTT(NTT(x)) = x, NTT(delta_0) = 1, convolution theorem.
 *   lde             brute-force closed form sum c_i (shift*w_M^k)^i (lg <= 6).
 *   naive_lde       periodicity NTT_N(c)[k mod N].
 *   gp / naive_gp   sequential big-int products.
 *   eval            big-int Horner.
 *   div             big-int synthetic division + q(X)(X-z) + a(z) = a(X).
 *   perm ratio      big-int restatement of the product formula.
 *   stage_no_rotate the reference driver's composition (bit-reverse, then per
 *                   step: ct_stage_indices gather, one call per chunk,
 *                   rotate_stage_output_index scatter) equals the NTT, for
 *                   schedules covering every factor rule (tests/test_segmented.py).
 */
#ifndef PL_ORACLE_H
#define PL_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef uint64_t or_fr[4];

/* field (PLONK/utils/mont/cuda/ff/mont_t.cuh:215-290, 396-419, 1117-1132) */
void or_add(uint64_t* o, const uint64_t* a, const uint64_t* b);
void or_sub(uint64_t* o, const uint64_t* a, const uint64_t* b);
void or_mul(uint64_t* o, const uint64_t* a, const uint64_t* b);
void or_neg(uint64_t* o, const uint64_t* a);
void or_pow(uint64_t* o, const uint64_t* a, const uint64_t* e4); /* 256-bit exponent */
void or_inv(uint64_t* o, const uint64_t* a);                    /* 0 -> 0 */
void or_to_mont(uint64_t* o, const uint64_t* a);
void or_from_mont(uint64_t* o, const uint64_t* a);
void or_one(uint64_t* o);
void or_from_u64(uint64_t* o, uint64_t v);  /* Montgomery form of a small integer */
int  or_is_canonical(const uint64_t* a);

/* vectorised field ops over n elements (element-wise ops, function.cuh:14-48) */
void or_vec_add(uint64_t* o, const uint64_t* a, const uint64_t* b, uint64_t n);
void or_vec_sub(uint64_t* o, const uint64_t* a, const uint64_t* b, uint64_t n);
void or_vec_mul(uint64_t* o, const uint64_t* a, const uint64_t* b, uint64_t n);
void or_vec_inv(uint64_t* o, const uint64_t* a, uint64_t n);
/* exp_mod, e = 0 -> one, e = 1 -> copy (cpu/mont_arithmetic.cpp:66-81) */
void or_vec_exp(uint64_t* o, const uint64_t* a, uint64_t n, uint32_t e);
/* gen_sequence o[i] = x^i by a running product (cuda/mont_arithmetic.cu:104-114) */
void or_sequence(uint64_t* o, const uint64_t* x, uint64_t n);

/* domain constants (parameters/bls12_381.h) */
void or_root_of_unity(uint64_t* o, int lg);       /* omega_{2^lg}, 0 <= lg <= 32 */
void or_domain_size_inverse(uint64_t* o, int lg); /* 2^-lg */
void or_group_gen(uint64_t* o);                   /* 7 */

/* bit reversal of the low `bits` bits */
uint64_t or_bitrev(uint64_t v, int bits);

/* seeded uniform Fr: SplitMix64, 4 draws per element, top bit cleared,
 * rejection of values >= r.  The accepted limbs are the Montgomery form. */
void or_fill_uniform(uint64_t* out, uint64_t n, uint64_t seed);
uint64_t or_seed(int op, int lg, int aux);

/* host threads used by the transforms (or_ntt, or_coset_ntt, or_lde,
 * or_naive_lde): default min(16, online CPUs) or $OR_THREADS; the results do
 * not depend on it */
void or_set_threads(int t);
int or_get_threads(void);

/* NTT, natural order in and out (Ntt::forward / Intt::forward,
 * PLONK/utils/function.cu:280-332).  In place. */
void or_ntt(uint64_t* data, int lg, int inverse);
/* coset NTT: out[k] = sum_i c_i (shift*w_N^k)^i, in place */
void or_coset_ntt(uint64_t* data, int lg, const uint64_t* shift);
/* inverse coset NTT: x_i = shift^-i iNTT(X)_i, in place (pinned by the round
 * trip with or_coset_ntt, whose column-0 role in or_lde the brute-force LDE
 * fixtures pin) */
void or_coset_intt(uint64_t* data, int lg, const uint64_t* shift);
/* our_lde (lde.cu:1267-1365): out[k] = sum_{i<N} c_i (shift*w_M^k)^i,
 * M = N << lg_blowup.  The reference's shift is 7^(2^lg_blowup). */
void or_lde(const uint64_t* coeffs, uint64_t* out, int lg_n, int lg_blowup, const uint64_t* shift);
/* primitive 3*2^lg-th root of unity 7^((r-1)/(3*2^lg)); its cube is w_{2^lg} */
void or_root_of_unity_3n(uint64_t* o, int lg);
/* blowup-3 LDE (no counterpart in the reference): out[k] = sum_i c_i
 * (shift*w_3N^k)^i, k < 3N, i.e. out[3j + t] = coset NTT with shift
 * shift*w_3N^t at j (pinned by brute force from Python integers) */
void or_lde3(const uint64_t* coeffs, uint64_t* out, int lg_n, const uint64_t* shift);
/* naive_lde (lde.cu:1367-1490): out[k] = NTT_N(c)[k mod N] */
void or_naive_lde(const uint64_t* coeffs, uint64_t* out, int lg_n, int lg_blowup);

/* accumulate_mul_poly (mont_arithmetic.cu:333-361): exclusive product,
 * z[0] = 1, z[i+1] = z[i]*p[i], in place */
void or_grand_product(uint64_t* data, uint64_t n);
/* naive_gp (miss_ops.cu:34-47): out[i] = prod_{j<=i} p[j] for i <= n-2,
 * out[n-1] = 0 */
void or_naive_gp(const uint64_t* in, uint64_t* out, uint64_t n);
/* evaluate (function.cu:170-181): sum c_i x^i */
void or_eval(const uint64_t* coeffs, uint64_t n, const uint64_t* x, uint64_t* y);
/* poly_div_poly (mont_arithmetic.cu:304-331, 531-542): quotient of a(X) by
 * (X + c); q[i] = sum_{j>i} a_j z^{j-i-1} with z = -c, q[n-1] = 0; in place.
 * Returns the remainder a(z) in rem if rem != 0. */
void or_div(uint64_t* data, uint64_t n, const uint64_t* c, uint64_t* rem);
/* naive_div (miss_ops.cu:49-87): long division by [c, d1] (d1 the leading
 * coefficient), n-1 quotient coefficients */
void or_naive_div(const uint64_t* in, uint64_t n, const uint64_t* c, const uint64_t* d1, uint64_t* q);

/* PLONK permutation ratio p_i = prod_j (w_j + beta*k_j*omega^i + gamma)
 *                              / prod_j (w_j + beta*s_j + gamma),  x/0 = 0
 * (term shape of zksnark_compute_permutation_constraint.cu:13-14).
 * w = {a,b,c}, s = {s1,s2,s3}, k = {k1,k2,k3} (Montgomery). */
void or_perm_ratio(const uint64_t* a, const uint64_t* b, const uint64_t* c,
                   const uint64_t* s1, const uint64_t* s2, const uint64_t* s3,
                   const uint64_t* beta, const uint64_t* gamma,
                   const uint64_t* k1, const uint64_t* k2, const uint64_t* k3,
                   int lg_n, uint64_t* ratio);

/* The segmented NTT's per-chunk step, forward direction: one call of
 * Ntt::stage_no_rotate(chunk, lg_domain, iterations, stage, chunk_id)
 * (function.cu:312-314 -> zksnark_ntt.cu:194-232, 440-454 -> ntt.cuh:642-665
 * -> CTkernel_no_rotate, ct_mixed_radix_wide.cuh:884-974 ->
 * _CT_NTT_without_rotate, ct_mixed_radix_wide.cuh:610-751), restated as its
 * thread program over the chunk's chunk_elems / 2 threads.  Returns 0, or -1
 * for a configuration the reference rejects (radix 9, sizes). */
int or_stage_no_rotate(uint64_t* chunk, uint64_t chunk_elems, int lg_domain, int iterations, int stage,
                       uint64_t chunk_id);
/* the driver's index algebra (ntt.cu:168-195) */
void or_ct_stage_indices(uint64_t tid, int stage, int iterations, uint64_t* idx0, uint64_t* idx1);
uint64_t or_rotate_stage_output_index(uint64_t global_idx, int stage, int iterations);

#ifdef __cplusplus
}
#endif
#endif
