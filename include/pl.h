/* libpl: BLS12-381 Fr prover operators for AMD Instinct MI355X (gfx950).
 *
 * C-ABI boundary of the library.  Every data buffer is device memory owned by
 * the caller, holding canonical Montgomery-form Fr elements of 32 bytes
 * (8 x u32 little-endian limbs == 4 x u64 limbs), 16-byte aligned.  The
 * library owns only domain tables (immutable after creation, shareable across
 * streams) and, when the caller passes no workspace, stream-ordered scratch.
 *
 * Calls are asynchronous on the given stream (NULL = the default stream) and
 * thread-safe for distinct streams.  Every entry point returns PL_OK (0) or a
 * negative pl_status; nothing in the library exits the process.
 *
 * The names in brackets are the reference interfaces each entry point
 * replaces (file:line in zlyber/ppopp26-artifact).
 */
#ifndef PL_H
#define PL_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ABI-identical to hipStream_t */
typedef struct ihipStream_t* pl_stream_t;

typedef struct pl_domain pl_domain;
typedef struct pl_lde_domain pl_lde_domain;

enum pl_status {
  PL_OK = 0,
  PL_ERR_ARG = -1,         /* bad size, alignment or null pointer */
  PL_ERR_HIP = -2,         /* a HIP runtime call failed */
  PL_ERR_UNSUPPORTED = -3, /* size outside the supported range */
  PL_ERR_NOMEM = -4,       /* device allocation failed */
  PL_ERR_NO_DEVICE = -5    /* no HIP device visible */
};

const char* pl_status_string(int status);
/* library version, major*10000 + minor*100 + patch */
int pl_version(void);
/* the HIP error behind the last PL_ERR_HIP / PL_ERR_NOMEM on this thread */
const char* pl_last_hip_error(void);

/* ---- NTT domains ------------------------------------------------------
 * [Ntt::Ntt / Intt::Intt, PLONK/utils/function.cu:278,328;
 *  params_zkp_cuda, PLONK/utils/zkp/cuda/zksnark_ntt.cu:17-39;
 *  NTTParameters, zksnark_ntt/parameters/parameters.cuh:142-222]
 * Twiddle tables for forward and inverse NTTs of size 2^lg, 1 <= lg <= 28,
 * resident in HBM on the current device. */
int pl_domain_create(int lg, pl_domain** out);
int pl_domain_destroy(pl_domain* d);
int pl_domain_lg(const pl_domain* d);
/* The reference's windowed power tables, the semi-public prefix of its
 * Params buffer (parameters.cuh:7-33, read by lde.cu:658-659): 4 x 16384
 * elements w_{2^28}^i, w_{2^28}^(i*2^14), 7^i, 7^(i*2^14). */
int pl_params_fill(void* d_out, pl_stream_t stream);
size_t pl_domain_bytes(const pl_domain* d);

/* ---- NTT / iNTT -------------------------------------------------------
 * [Ntt::forward / Ntt::forward_, function.cu:280-286;
 *  Intt::forward / Intt::forward_, function.cu:328-336;
 *  ntt_zkp_cuda, zksnark_ntt.cu:342-381]
 * X[k] = sum_j x[j] w^(jk), w = 7^((r-1)/2^lg), natural order in and out.
 * inverse != 0: w^-1 and a final factor 2^-lg.  d_in may equal d_out.
 * d_work: NULL, or pl_ntt_work_bytes(lg, 1) bytes of device scratch. */
int pl_ntt(const void* d_in, void* d_out, int lg, int inverse, const pl_domain* d,
           void* d_work, pl_stream_t stream);
/* npoly contiguous polynomials of size 2^lg in one launch sequence */
int pl_ntt_batch(const void* d_in, void* d_out, int lg, uint32_t npoly, int inverse,
                 const pl_domain* d, void* d_work, pl_stream_t stream);
/* workspace for pl_ntt_batch: npoly * 2^lg elements for lg >= 6 (any pass
 * split of such a size may need it), 0 below */
size_t pl_ntt_work_bytes(int lg, uint32_t npoly);

/* Pass-level access to the same transform (the intermediate layout is this
 * library's own, and intermediate values are 32-byte packed balanced-digit
 * words congruent to the field elements, not canonical Montgomery values;
 * the reference's per-chunk step with its own layout is
 * pl_ntt_stage_no_rotate below).  Running passes
 * 0 .. pl_domain_num_passes(d)-1 in order, each reading the previous pass's
 * output, equals pl_ntt_batch.  Intermediate passes may run in place; the
 * last pass writes natural order and must not (d_in != d_out).
 * pl_ntt_num_passes(lg) is the pass count of a domain created now (the
 * split in use, see pl_plan_set); a domain keeps the split it was built
 * with. */
int pl_ntt_num_passes(int lg);
int pl_domain_num_passes(const pl_domain* d);
int pl_ntt_pass(const void* d_in, void* d_out, int lg, uint32_t npoly, int pass, int inverse,
                const pl_domain* d, pl_stream_t stream);

/* The segmented NTT's per-chunk step with the reference's own semantics
 * [Ntt::stage_no_rotate, PLONK/utils/function.cu:312-314;
 *  _CT_NTT_without_rotate, zksnark_ntt/ntt_kernel/ct_mixed_radix_wide.cuh:610-751,
 *  launched by CTkernel_no_rotate, ct_mixed_radix_wide.cuh:884-974].
 * d_chunk holds chunk_elems (a power of two >= 2^iterations) elements
 * gathered with the driver's ct_stage_indices (ntt.cu:168-178) as pairs
 * (2t, 2t+1); the call runs DIT stages [stage, stage + iterations) of a
 * forward NTT of size 2^lg_domain (lg_domain <= 28) in place, the result left
 * in gathered order; twiddles use the global thread chunk_id*chunk_elems/2 + t.
 * iterations in 1..8 or 10 (radix max(6, iterations) in {6,7,8,10}; 9 is
 * rejected as by the reference).  d_params: a Params buffer (pl_params_fill). */
int pl_ntt_stage_no_rotate(void* d_chunk, uint64_t chunk_elems, int lg_domain, int iterations, int stage,
                           uint64_t chunk_id, const void* d_params, pl_stream_t stream);

/* The chunked flow's index algebra [ct_stage_indices, ntt.cu:168-178;
 * rotate_stage_output_index, ntt.cu:189-195] (host functions): thread tid
 * of a step (stage, iterations) owns the global pair (*idx0, *idx1 =
 * *idx0 + 2^stage); a step's output element g belongs at g with bits
 * [stage, stage + iterations) rotated right by one. */
void pl_ct_stage_indices(uint64_t tid, int stage, int iterations, uint64_t* idx0, uint64_t* idx1);
uint64_t pl_rotate_stage_output_index(uint64_t global_idx, int stage, int iterations);

/* The chunk permutations as device kernels (chunk = chunk_elems elements,
 * a power of two; buffers distinct):
 *   pl_stage_gather      chunk[2t+b] = src[idx_b(chunk_id*chunk_elems/2 + t)]
 *                        [permute_stage_input_chunk_range, ntt.cu:335-354]
 *   pl_stage_scatter     dst[rotate(idx_b(...)) - (chunk_local ? chunk_id*chunk_elems : 0)] = chunk[2t+b]
 *                        [scatter_stage_output_chunk_range, ntt.cu:593-614;
 *                         scatter_stage_kernel, ntt.cu:647-700]
 *   pl_stage_transition  dst[2t+b] = src[rotate_prev^-1(idx_b^next(...)) - chunk_id*chunk_elems]:
 *                        a chunk left in one step's rotated order regathered
 *                        for the next step [transition_stage_kernel, ntt.cu:702-756]
 *   pl_bit_reverse       out[rev_lg(i)] = in[i], in != out [the host bit
 *                        reversal of our_ntt: calculate_map1 +
 *                        bit_rev_step1_parallel, function.cu:503-509,
 *                        function.cuh:221-244] */
int pl_stage_gather(const void* d_src, void* d_chunk, uint64_t chunk_elems, uint64_t chunk_id, int stage,
                    int iterations, pl_stream_t stream);
int pl_stage_scatter(const void* d_chunk, void* d_dst, uint64_t chunk_elems, uint64_t chunk_id, int stage,
                     int iterations, int chunk_local, pl_stream_t stream);
int pl_stage_transition(const void* d_src, void* d_dst, uint64_t chunk_elems, uint64_t chunk_id, int prev_stage,
                        int prev_iterations, int next_stage, int next_iterations, pl_stream_t stream);
int pl_bit_reverse(const void* d_in, void* d_out, int lg, pl_stream_t stream);

/* The reference's segmented NTT flow, device-resident [our_ntt,
 * ntt.cu:979-1037 with run_first_ntt_pass, ntt.cu:801-859, and
 * run_second_ntt_pass_async_scatter, ntt.cu:861-977]: bit reversal, pass 1
 * over chunks of 2^(lg_i + lg_k) consecutive elements (stage_no_rotate of
 * lg_i/2 and lg_i - lg_i/2 stages with the transition between them and a
 * chunk-local scatter), pass 2 per chunk (gather for stage lg_i, one step of
 * lg_n - lg_i stages, scatter).  Forward NTT, natural order in and out, in
 * place; every permutation is a device kernel.  The same split rules as the
 * reference: lg_i/2, lg_i - lg_i/2 and lg_n - lg_i in 1..8 or 10.
 * d_params: pl_params_fill; d_work: NULL or pl_ntt_segmented_work_bytes. */
size_t pl_ntt_segmented_work_bytes(int lg_n, int lg_i, int lg_k);
int pl_ntt_segmented(void* d_inout, int lg_n, int lg_i, int lg_k, const void* d_params, void* d_work,
                     pl_stream_t stream);

/* Host-to-host NTT [our_ntt's scope, ntt.cu:979-1037]: npoly polynomials of
 * 2^lg elements in host memory (pinned for overlap) in, results out
 * (h_out may equal h_in).  Two device slots and two copy streams beside
 * `stream`: each polynomial's first pass runs on column groups as they
 * arrive (2^lg_k groups, each a strided 2-D copy), its last pass on line
 * groups whose natural-order outputs leave at once, and the next
 * polynomial's copy-in overlaps the current one's transform and copy-out.
 * lg_k < 0: the count in use for this size (pl_ntt_host_groups).  Blocks
 * until h_out holds every result.  d_work: NULL or
 * pl_ntt_host_work_bytes(lg, npoly) bytes. */
size_t pl_ntt_host_work_bytes(int lg, uint32_t npoly);
int pl_ntt_host(const void* h_in, void* h_out, int lg, uint32_t npoly, int inverse, const pl_domain* d, int lg_k,
                void* d_work, pl_stream_t stream);
/* The copy-group count (lg_k) per size for pl_ntt_host's lg_k < 0 [the
 * reference's segment count k of our_ntt, swept by hand with NTT
 * --bench-sweep, ntt.cu:1494-1500]: the autotuner's choice (PL_TUNE_HOST) or
 * a cache file's, else 2.  _set: lg_k < 0 restores the default; at most 20. */
int pl_ntt_host_groups(int lg);
int pl_ntt_host_groups_set(int lg, int lg_k);

/* ---- coset low-degree extension ---------------------------------------
 * [our_lde, lde.cu:1267-1365 (lde_spread_powers_kernel lde.cu:623-648,
 *  make_lde_plan lde.cu:288-302)]
 * N = 2^lg_n coefficients c (natural order) -> M = N << lg_blowup values
 *   out[k] = sum_{i<N} c_i (shift * w_M^k)^i,  k < M, natural order.
 * h_shift: host pointer to one Montgomery element, or NULL for the
 * reference's shift 7^(2^lg_blowup).  Computed as 2^lg_blowup coset NTTs of
 * size N; the coset powers ride on the first pass, the interleave on the
 * last.  1 <= lg_blowup <= 4, lg_n + lg_blowup <= 28.
 * d_out may equal d_coeffs when that buffer holds M elements.
 * d_work: NULL or pl_lde_work_bytes(lg_n, lg_blowup) bytes. */
int pl_lde_domain_create(int lg_n, int lg_blowup, const void* h_shift, pl_lde_domain** out);
int pl_lde_domain_destroy(pl_lde_domain* d);
size_t pl_lde_domain_bytes(const pl_lde_domain* d);
size_t pl_lde_work_bytes(int lg_n, int lg_blowup);
int pl_lde(const void* d_coeffs, void* d_out, const pl_lde_domain* d, void* d_work, pl_stream_t stream);
/* Blowup 3 (no counterpart in the reference, whose our_lde takes
 * power-of-two blowups; run as three coset NTTs of size N with
 * shifts g w_3N^t): M = 3N outputs out[k] = sum_i c_i (g w_3N^k)^i, natural
 * order, w_3N = 7^((r-1)/(3N)) (its cube is w_N), g = 7 when h_shift is
 * NULL; lg_n <= 26.  Run by pl_lde on such a domain: three coset NTTs of
 * size N with shifts g w_3N^t side by side, then the interleave
 * out[3 j + t].  pl_lde_domain_work_bytes gives the workspace of any LDE
 * domain (6N elements for blowup 3), pl_lde_domain_blowup its blowup. */
int pl_lde3_domain_create(int lg_n, const void* h_shift, pl_lde_domain** out);
size_t pl_lde_domain_work_bytes(const pl_lde_domain* d);
int pl_lde_domain_blowup(const pl_lde_domain* d);

/* ---- coset NTT / iNTT ----------------------------------------------------
 * [Ntt_coset / Intt_coset, PLONK/utils/function.cuh:116-142,
 *  function.cu:354-395; ntt_zkp_coset_*, zksnark_ntt.cu:521-908]
 * Over the shifted domain g<w_N>, N = 2^lg, g = 7 (h_shift NULL) or a
 * caller-given Montgomery element:
 *   forward  X[k] = sum_i x_i (g w_N^k)^i
 *   inverse  x_i  = g^-i (1/N) sum_k X_k w_N^-ik
 * pl_coset_ntt runs the whole transform (d_in may equal d_out; d_work NULL
 * or pl_coset_work_bytes(lg)).  pl_coset_scale is the coset-power step alone
 * on chunk chunk_id of a natural-order vector (chunk[i] *= g^(+-(chunk_id *
 * chunk_elems + i)), in place; the reference's init).  pl_coset_ntt_step
 * splits the transform into the reference's three calls (forward1/2/3) over
 * the whole vector: step 1 the first pass with the coset powers (d_in ->
 * d_out, may alias), step 2 the middle passes (in place on d_in), step 3 the
 * last pass into natural order (d_in -> d_out, distinct; inverse: times
 * g^-i); steps 1..3 in order equal pl_coset_ntt. */
typedef struct pl_coset_domain pl_coset_domain;
int pl_coset_domain_create(int lg, const void* h_shift, pl_coset_domain** out);
int pl_coset_domain_destroy(pl_coset_domain* d);
size_t pl_coset_domain_bytes(const pl_coset_domain* d);
size_t pl_coset_work_bytes(int lg);
int pl_coset_ntt(const void* d_in, void* d_out, int inverse, const pl_coset_domain* d, void* d_work,
                 pl_stream_t stream);
int pl_coset_scale(void* d_chunk, uint64_t chunk_elems, uint64_t chunk_id, int inverse, const pl_coset_domain* d,
                   pl_stream_t stream);
int pl_coset_ntt_step(const void* d_in, void* d_out, int step, int inverse, const pl_coset_domain* d,
                      pl_stream_t stream);

/* ---- scans -------------------------------------------------------------
 * All three are work-efficient block scans (~3 Montgomery products per
 * element).  d_work: NULL or pl_scan_work_bytes(n) bytes of device scratch.
 *
 * Grand product [accumulate_mul_poly / accumulate_mul_poly_,
 * PLONK/utils/function.cu:256-262; mont_arithmetic.cu:181-201, 333-361, 544-556]:
 * in place, d_inout[i] <- prod_{j<i} p[j] (exclusive; element 0 becomes
 * Mont(1)).  d_incl_out: NULL, or receives the inclusive products
 * prod_{j<=i} p[j] (what accumulate_mul_poly_ leaves in its output). */
size_t pl_scan_work_bytes(uint64_t n);
int pl_grand_product(void* d_inout, uint64_t n, void* d_incl_out, void* d_work, pl_stream_t stream);
/* The same grand product with no allocation and no workspace argument (no
 * counterpart in the reference): for callers that cannot allocate, such as
 * code under HIP graph capture or a request loop that must not call
 * hipMalloc.  Its block partials live in a 4 MiB device array the library
 * holds per GPU until the process ends; n <= 2^28 (PL_ERR_UNSUPPORTED
 * beyond).  Calls on one GPU share that array: calls that overlap on the
 * device (different streams, not ordered by events) race. */
int pl_grand_product_noalloc(void* d_inout, uint64_t n, void* d_incl_out, pl_stream_t stream);

/* Evaluation [evaluate, function.cu:170-181; mont_arithmetic.cu:104-163]:
 * *d_y = sum_{i<n} poly[i] * x^i, x a device element (the reference passes
 * the point as a device buffer).  n == 0 gives 0. */
int pl_eval(const void* d_poly, uint64_t n, const void* d_x, void* d_y, void* d_work, pl_stream_t stream);

/* Division by (X + c) [poly_div_poly, function.cu:183-185;
 * mont_arithmetic.cu:165-179, 203-212, 304-331, 531-542]:
 * in place, d_inout[i] <- q[i] = sum_{j>i} a_j z^(j-i-1), z = -c, and
 * q[n-1] = 0.  d_rem: NULL, or receives the remainder a(z) (the reference
 * discards it).  c is a device element and is not modified. */
int pl_div_linear(void* d_inout, uint64_t n, const void* d_c, void* d_rem, void* d_work, pl_stream_t stream);

/* ---- pass-split autotuner ----------------------------------------------------
 * (replaces the reference's manual segment sweep, NTT/LDE --bench-sweep,
 *  ntt.cu:1494-1500, lde.cu:1834-1846)
 * A size-2^lg transform runs as P passes of 2^m[p]-point columns (3 <= m <= 6,
 * sum = lg); every split gives the same exact result.  pl_autotune times the
 * candidate splits of each lg in [lg_lo, lg_hi] on the current device (a
 * batch of max(1, 2^22 / N) polynomials per call), keeps the fastest and
 * records the choices in cache_path ("<lg> <P> <m0..> <ms>" lines); sizes
 * already in the cache are not re-measured unless PL_TUNE_REMEASURE.  The
 * chosen splits apply to domains created afterwards (process-wide).
 * pl_plan_get reports the split in use (P = 0 below lg 4: one tiny kernel);
 * pl_plan_set sets one (P = 0 restores the default); pl_plan_load applies a
 * cache file and returns the number of entries it applied (splits and
 * copy-group counts); pl_plan_candidates writes
 * up to max_cand splits (8 ints each, zero-padded) and returns their count.
 * PL_TUNE_HOST: also time pl_ntt_host (one polynomial, pinned host buffers)
 * for each lg in [max(lg_lo, 12), min(lg_hi, 26)] over lg_k = 0..6 and keep
 * the fastest as that size's copy-group count ("h <lg> <lg_k> <ms>" lines in
 * the cache; pl_plan_load applies them too). */
enum pl_tune_flags { PL_TUNE_REMEASURE = 1, PL_TUNE_HOST = 2 };
int pl_autotune(int lg_lo, int lg_hi, const char* cache_path, int reps, int flags);
int pl_plan_get(int lg, int* m, int* P);
int pl_plan_set(int lg, const int* m, int P);
int pl_plan_load(const char* path);
int pl_plan_candidates(int lg, int* out, int max_cand);

/* ---- baselines ------------------------------------------------------------
 * [standard4_ntt, ntt.cu:1039-1138; naive_lde, lde.cu:1367-1490]
 * Workload baselines for the CLIs: bit reversal plus one radix-2 launch per
 * stage.  pl_ntt_naive: same result as pl_ntt (forward), d_in != d_out.
 * pl_lde_naive: the reference's naive_lde result, out[k] = NTT_N(c)[k mod N]
 * for k < N << lg_blowup (periodic, no coset). */
int pl_ntt_naive(const void* d_in, void* d_out, int lg, const pl_domain* d, pl_stream_t stream);
int pl_lde_naive(const void* d_coeffs, void* d_out, int lg_n, int lg_blowup, const pl_domain* d,
                 pl_stream_t stream);

/* ---- element-wise operators ---------------------------------------------
 * [add_mod / sub_mod / mul_mod / div_mod and the _scalar forms, neg_mod,
 *  inv_mod, to_mont, to_base: PLONK/utils/function.cuh:12-48, function.cu:5-152]
 * out[i] = a[i] op b[i]; for the *_SCALAR ops b is one device element.
 * DIV and INV follow the reference's x/0 = 0 (mont_t.cuh:1117-1126) and use
 * one field inversion per call (Montgomery's trick, block-hierarchical).
 * out may alias a or b.  d_work: NULL or pl_vec_work_bytes(n) bytes. */
enum pl_vec_op_code {
  PL_VEC_ADD = 0, PL_VEC_SUB = 1, PL_VEC_MUL = 2, PL_VEC_DIV = 3,
  PL_VEC_ADD_SCALAR = 4, PL_VEC_SUB_SCALAR = 5, PL_VEC_MUL_SCALAR = 6, PL_VEC_DIV_SCALAR = 7,
  PL_VEC_NEG = 8, PL_VEC_INV = 9, PL_VEC_TO_MONT = 10, PL_VEC_FROM_MONT = 11
};
size_t pl_vec_work_bytes(uint64_t n);
int pl_vec_op(int op, const void* d_a, const void* d_b, void* d_out, uint64_t n, void* d_work,
              pl_stream_t stream);

/* [exp_mod: PLONK/utils/function.cuh:32, function.cu:113-120,
 *  mont/cuda/mont_arithmetic.cu:258-273, 450-463]
 * out[i] = in[i]^exp; exp = 0 writes one, exp = 1 copies.  The reference
 * traps on a negative exponent; here that is PL_ERR_ARG.  out may alias in. */
int pl_exp_mod(const void* d_in, void* d_out, uint64_t n, int exp, pl_stream_t stream);
/* [gen_sequence / gen_sequence_: function.cuh:46-47, function.cu:154-160,
 *  poly_eval_kernel mont_arithmetic.cu:104-114]
 * out[i] = x^i for i < n (out[0] = one, out[1] = x); n <= 2^28; out must not
 * overlap x. */
int pl_gen_sequence(uint64_t n, const void* d_x, void* d_out, pl_stream_t stream);
/* [repeat_to_poly / repeat_to_poly_: function.cuh:49-50, function.cu:162-168,
 *  repeat_kernel mont_arithmetic.cu:214-220]
 * out[i] = *x for i < n; out must not overlap x. */
int pl_repeat_to_poly(const void* d_x, void* d_out, uint64_t n, pl_stream_t stream);

/* ---- PLONK permutation ratio --------------------------------------------
 * (the quotient term shape of zksnark_compute_permutation_constraint.cu:13-14,
 *  built from element-wise calls in the reference)
 * ratio[i] = prod_j (w_j[i] + beta*k_j*w^i + gamma) / prod_j (w_j[i] + beta*s_j[i] + gamma)
 * for j = 0..2, w = omega_{2^lg} from the domain d (d->lg == lg), x/0 = 0.
 * d_w, d_s: three device columns each; beta, gamma device elements; h_k three
 * host Montgomery constants (k1, k2, k3).  d_work: NULL or
 * pl_vec_work_bytes(2^lg) bytes.  pl_perm_grand_product writes the
 * exclusive grand product of the ratio to d_z (the PLONK z column: z[0] =
 * Mont(1), z[i] = prod_{j<i} ratio[j]) without forming the ratio vector:
 * z[i] = prod_{j<i} num[j] * prod_{j>=i} den[j] / prod_j den[j], one field
 * inversion per call; same arguments, same d_work size. */
int pl_perm_ratio(const void* const* d_w, const void* const* d_s, const void* d_beta, const void* d_gamma,
                  const void* h_k, int lg, const pl_domain* d, void* d_ratio, void* d_work, pl_stream_t stream);
int pl_perm_grand_product(const void* const* d_w, const void* const* d_s, const void* d_beta,
                          const void* d_gamma, const void* h_k, int lg, const pl_domain* d, void* d_z,
                          void* d_work, pl_stream_t stream);

/* ---- batched form: the uneven-size polynomials of one proof ---------------
 * (no single counterpart: the reference runs one our_ntt / our_lde /
 *  evaluate / poly_div_poly call per polynomial, ntt.cu:979-1037,
 *  lde.cu:600-688, mont_arithmetic.cu:104-163, zksnark_poly_div.cu)
 * A batch is planned once over caller-owned device buffers: the transforms'
 * passes are grouped by depth and radix into ragged launches (one launch
 * covers the tiles of many polynomials), evaluation + division pairs are one
 * suffix scan each on a forked stream, and with PL_BATCH_GRAPH the whole
 * sequence is captured as one hipGraph.  pl_batch_run allocates nothing and
 * may be repeated; the buffers named by the items must stay valid until
 * pl_batch_destroy.  Results are bit-identical to the per-call entry points:
 *   NTT / INTT  out = pl_ntt(in, lg, inverse)          (out may alias in)
 *   LDE         out = pl_lde(in) over 7^(2^lg_blowup) * <w_M>, M = 2^(lg + lg_blowup)
 *   EVAL_DIV    y = p(z); out = quotient of (p(X) - p(z)) by (X - z), out[n-1] = 0,
 *               p = in (n = 2^lg coefficients), z a device element; out != in
 * pl_batch_run_plain runs the same items as one launch sequence per
 * polynomial on `stream` (the per-call entry points' kernels), for
 * comparison.  Both read the items' inputs and points when they run, so the
 * caller may rewrite them (a new challenge z per proof) between runs.  Runs
 * of one batch share its scratch: the caller orders them (one stream, or
 * events).  pl_batch_destroy waits for the batch's last run and its own
 * streams, not for the rest of the device. */
enum pl_batch_op { PL_BATCH_NTT = 0, PL_BATCH_INTT = 1, PL_BATCH_LDE = 2, PL_BATCH_EVAL_DIV = 3 };
enum pl_batch_flags { PL_BATCH_GRAPH = 1 };
typedef struct pl_batch_item {
  int op;         /* pl_batch_op */
  int lg;         /* log2 of the input length, 0..28 (LDE: lg + lg_blowup <= 28) */
  int lg_blowup;  /* LDE only, 1..4 */
  const void* d_in;
  void* d_out;
  const void* d_z; /* EVAL_DIV: the point */
  void* d_y;       /* EVAL_DIV: p(z) */
} pl_batch_item;
typedef struct pl_batch pl_batch;
int pl_batch_create(const pl_batch_item* items, int n_items, int flags, pl_batch** out);
int pl_batch_run(pl_batch* b, pl_stream_t stream);
int pl_batch_run_plain(pl_batch* b, pl_stream_t stream);
/* launches per pl_batch_run (graph nodes when captured) and device bytes the
 * plan owns (scratch + launch tables; domains are shared library-wide) */
int pl_batch_info(const pl_batch* b, int* n_launches, size_t* bytes);
void pl_batch_destroy(pl_batch* b);

/* ---- served form: a request loop for proof jobs ---------------------------
 * (no counterpart in the reference, which runs one operator call at a time)
 * A job is one proof's eight columns e_0..e_7 of N = 2^lg evaluations (device
 * buffers the caller keeps valid until the job is done), beta, gamma and an
 * opening point zeta (device elements).  It computes, bit-identical to the
 * per-call entry points:
 *   coeffs[j] = iNTT(e_j)                                     8 x N elements
 *   lde[j]    = LDE(coeffs[j]) over 7^(2^lg_blowup) <w_M>     8 x M, M = N 2^lg_blowup
 *   z         = pl_perm_grand_product(w = e_0..e_2, s = e_3..e_5, beta, gamma,
 *               k = (1, 7, 13))                               N
 *   y[j]      = coeffs[j](zeta)                               8
 *   q[j]      = (coeffs[j] - y[j]) / (X - zeta), q[j][N-1] = 0  8 x N
 * into a server-owned block that stays in HBM until pl_server_release.
 * The server takes an arena of hbm_bytes at creation and admits queued jobs
 * in arrival order when their output and work blocks fit in it.  Batched
 * mode merges consecutive small jobs into one group, runs a group's
 * transforms as ragged launches and overlaps up to `slots` groups on the
 * device, at most one of them holding a large job (transform work
 * 8 N (1 + 2^lg_blowup) >= large_work; small jobs may start ahead of a large
 * one waiting for that); plain
 * mode runs one job at a time as per-call launch sequences.
 * pl_server_wait blocks until the job is done and reports the server-clock
 * times (seconds since creation) of its arrival (submit), admission and
 * completion.  A job that can never fit fails with PL_ERR_NOMEM. */
enum pl_serve_mode { PL_SERVE_BATCHED = 0, PL_SERVE_PLAIN = 1 };
typedef struct pl_server_config {
  size_t hbm_bytes;      /* arena size; 0 = free device memory less 32 GiB for domain tables */
  int slots;             /* groups running at once, 1..8; 0 = 2 (plain: 1) */
  int mode;              /* pl_serve_mode */
  int max_group_jobs;    /* jobs merged into one group; 0 = 16 */
  uint64_t group_elems;  /* merge while a group's 8 N sum stays <= this; 0 = 2^22 */
  uint64_t large_work;   /* a job is large when 8 N (1 + 2^lg_blowup) >= this; 0 = 2^25 */
} pl_server_config;
typedef struct pl_job_desc {
  int lg;         /* 1..27 */
  int lg_blowup;  /* 1..4, lg + lg_blowup <= 28 */
  const void* d_evals[8];
  const void* d_beta;
  const void* d_gamma;
  const void* d_zeta;
} pl_job_desc;
typedef struct pl_job_result {
  int status;
  double arrival_s, admit_s, done_s;
  const void* d_coeffs; /* [8][N] */
  const void* d_lde;    /* [8][M] */
  const void* d_z;      /* [N] */
  const void* d_y;      /* [8] */
  const void* d_q;      /* [8][N] */
} pl_job_result;
typedef struct pl_server_stats {
  uint64_t submitted, completed, groups, queued;
  size_t arena_bytes, arena_peak, arena_used, domain_bytes;
} pl_server_stats;
typedef struct pl_server pl_server;
int pl_server_create(const pl_server_config* cfg, pl_server** out);
int pl_server_submit(pl_server* s, const pl_job_desc* job, uint64_t* job_id);
int pl_server_wait(pl_server* s, uint64_t job_id, pl_job_result* result);
int pl_server_release(pl_server* s, uint64_t job_id);
int pl_server_get_stats(pl_server* s, pl_server_stats* stats);
double pl_server_now(const pl_server* s);
/* output block bytes of one job (0 for unsupported sizes) */
size_t pl_job_bytes(int lg, int lg_blowup);
void pl_server_destroy(pl_server* s);

/* ---- synthetic inputs and digests -----------------------------------------
 * (test and benchmark infrastructure; no counterpart in the reference)
 * pl_fill_uniform writes n elements uniform in [0, r) drawn from the seeded
 * SplitMix64 stream (4 outputs per candidate as little-endian limbs, top bit
 * cleared, rejected while >= r): the same values as the host generator on
 * the same seed.  d_work: NULL or pl_uniform_work_bytes(n) bytes.
 * pl_digest writes to d_out (8 bytes) sum_k (w_k ^ k*0x9E3779B97F4A7C15) *
 * (2k+1) mod 2^64 over the 64-bit words w_k of n elements. */
size_t pl_uniform_work_bytes(uint64_t n);
int pl_fill_uniform(void* d_out, uint64_t n, uint64_t seed, void* d_work, pl_stream_t stream);
int pl_digest(const void* d_data, uint64_t n, void* d_out, pl_stream_t stream);
/* Measured peak of 32x32->64-bit multiply-adds (v_mad_u64_u32) on the
 * current device, in 10^12 per second: independent chains in every lane of
 * a grid filling every CU (the roofline the Montgomery kernels are held to).
 * Synchronizes the stream. */
int pl_probe_mad_rate(double* tmad_per_s, pl_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif
