/*
 * mzk_gpu.h -- C ABI of the MI355X (gfx950) prover hot path: BLS12-381 G1
 * multi-scalar multiplication behind KZG commitments and scalar-field NTTs behind
 * EvaluationDomain.
 *
 * Conventions (all entry points):
 *   - return MZK_OK (0) or a non-zero status; never abort; mzk_last_error() holds
 *     a thread-local message for the last failure on the calling thread;
 *   - fr elements are 4 x u64 little-endian Montgomery limbs (R = 2^256), the
 *     memory image of the reference's `Fq` / blst_fr (curves/src/bls12_381/fq.rs:27-29);
 *   - G1 affine points are 12 x u64 (x[6] || y[6], Montgomery, R = 2^384; identity
 *     all zero), the image of `G1Affine` / blst_p1_affine (g1.rs:32-34);
 *   - G1 projective results are 18 x u64 Jacobian (X || Y || Z, identity Z = 0),
 *     the image of `G1Projective` / blst_p1 (g1.rs:493-495);
 *   - `_dev` entry points take device pointers (HBM-resident data) and run on the
 *     context's stream; host-pointer entry points copy in, compute and copy out;
 *   - calls on one context are serialised by the context; independent threads may
 *     use independent contexts concurrently.
 */
#ifndef MZK_GPU_H
#define MZK_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  MZK_OK = 0,
  MZK_ERR_INVALID = 1,     /* bad argument: length mismatch, size out of range, null pointer */
  MZK_ERR_HIP = 2,         /* HIP runtime failure */
  MZK_ERR_NOMEM = 3,       /* device allocation failed or a budget was exceeded */
  MZK_ERR_UNSUPPORTED = 4  /* valid request outside what this build implements */
};

/* Message for the last failure on this thread ("" if none). */
const char *mzk_last_error(void);
/* Library ABI version (major * 100 + minor). */
int mzk_version(void);

/* ------------------------------------------------------------------ context */
typedef struct mzk_ctx mzk_ctx;
/* Binds a context to HIP device `device`; owns one non-blocking stream and the
 * scratch memory its calls need. */
int mzk_ctx_create(int device, mzk_ctx **out);
void mzk_ctx_destroy(mzk_ctx *ctx);
/* The context's hipStream_t, for callers that record events or order work. */
void *mzk_ctx_stream(mzk_ctx *ctx);
int mzk_ctx_synchronize(mzk_ctx *ctx);
/* Free and total device memory of the context's device (hipMemGetInfo), for callers that
 * admit work by HBM budget.  No reference counterpart. */
int mzk_ctx_mem_info(mzk_ctx *ctx, size_t *free_bytes, size_t *total_bytes);

/* ------------------------------------------------------------------ MSM
 * Replaces curves/src/msm.rs:460 `msm_best(coeffs: &[C::Scalar], bases: &[C]) -> C::Curve`
 * (panics on length mismatch there; here both slices share the length n).
 * Host pointers; out = sum_i coeffs[i] * bases[i] as Jacobian.  n = 0 gives the identity. */
int mzk_msm_best(mzk_ctx *ctx, const uint64_t *coeffs, const uint64_t *bases, size_t n,
                 uint64_t out[18]);
/* Replaces proofs/src/poly/kzg/msm.rs:174 `msm_specific(coeffs, bases) -> C::Curve`:
 * zero scalars contribute nothing; all-zero (or empty) input gives the identity. */
int mzk_msm_specific(mzk_ctx *ctx, const uint64_t *coeffs, const uint64_t *bases, size_t n,
                     uint64_t out[18]);
/* Replaces curves/src/msm.rs:430 `msm_parallel(coeffs, bases) -> C::Curve` (the reference's
 * thread split of msm_serial chunks; the sum is the same unique point as msm_best). */
int mzk_msm_parallel(mzk_ctx *ctx, const uint64_t *coeffs, const uint64_t *bases, size_t n,
                     uint64_t out[18]);
/* Replaces curves/src/msm.rs:338 `msm_serial(coeffs, bases, acc: &mut C::Curve)`, acc
 * Jacobian in and out.  As the reference does, the incoming acc is doubled c times per
 * window before the sum is added (msm.rs:364-367): acc <- [2^(c W)] acc + sum_i coeffs[i]
 * bases[i], c = 1 (n < 4), 3 (n < 32) or ceil(ln n), W = 8 (highest non-zero byte of the
 * canonical scalars) / c + 1; all-zero scalars leave acc unchanged.  From the identity
 * (how msm_parallel calls it) acc becomes the MSM. */
int mzk_msm_serial(mzk_ctx *ctx, const uint64_t *coeffs, const uint64_t *bases, size_t n,
                   uint64_t acc[18]);
/* Same result with device-resident scalars and bases. */
int mzk_msm_dev(mzk_ctx *ctx, const void *d_coeffs, const void *d_bases, size_t n,
                uint64_t out[18]);

/* Jacobian -> affine on the host (G1Projective::to_affine; g1.rs:725-739 normalises in
 * batches with the same result): x = X/Z^2, y = Y/Z^3; the identity maps to all zeros. */
int mzk_g1_jacobian_to_affine(const uint64_t jac[18], uint64_t aff[12]);

/* ------------------------------------------------------------------ resident SRS
 * Replaces the bases side of proofs/src/poly/kzg/params.rs:22-37 (`ParamsKZG::g`) and
 * `bases::<B>()` (params.rs:83-93): the bases stay in HBM across commits. */
typedef struct mzk_srs mzk_srs;
/* Upload n affine bases (host pointer). */
int mzk_srs_create(mzk_ctx *ctx, const uint64_t *bases, size_t n, mzk_srs **out);
/* g_i = [tau^i] G1 for i < n, computed on the device (params.rs:154-172, unsafe_setup). */
int mzk_srs_generate(mzk_ctx *ctx, const uint64_t tau[4], size_t n, mzk_srs **out);
/* Bases [s_i] G1 for n device scalars s_i: the Lagrange basis of unsafe_setup
 * (params.rs:174-186, g_lagrange_i = [L_i(tau)] G1) takes s_i = L_i(tau). */
int mzk_srs_from_scalars_dev(mzk_ctx *ctx, const void *d_scalars, size_t n, mzk_srs **out);
/* g_to_lagrange (proofs/src/utils/arithmetic.rs:34-54): the Lagrange bases of a
 * monomial SRS without its trapdoor -- the inverse DFT over G1 of g[0, 2^k) with
 * omega = ROOT^(2^(32-k)), times 1/2^k, normalised to affine.  A new SRS of 2^k bases. */
int mzk_srs_to_lagrange(mzk_ctx *ctx, const mzk_srs *g, uint32_t k, mzk_srs **out);
/* suffix_sum (proofs/src/poly/kzg/params.rs:47-57): out_i = sum_(j >= i) in_j, the
 * g_lagrange_delta bases (and, applied to those, g_lagrange_double_delta). */
int mzk_srs_suffix_sum(mzk_ctx *ctx, const mzk_srs *in, mzk_srs **out);
/* G1 point encodings of the SRS file (proofs/src/utils/helpers.rs:16-28 SerdeFormat,
 * same order). */
enum {
  MZK_SERDE_PROCESSED = 0,          /* 48 B compressed (ZCash flags 0x80/0x40/0x20) */
  MZK_SERDE_RAW_BYTES = 1,          /* 96 B uncompressed big-endian x || y */
  MZK_SERDE_RAW_BYTES_UNCHECKED = 2 /* the same bytes; on-curve check only */
};
/* The point arrays of ParamsKZG::read_custom (params.rs:286-336): count encoded G1
 * points (48 or 96 bytes each, host memory) decoded on the device into a new SRS.
 * Processed decompresses (from_compressed, g1.rs:367-370) and RawBytes parses
 * (from_uncompressed, read_raw g1.rs:472-484); both check the curve equation and
 * membership of the prime-order subgroup; RawBytesUnchecked (read_raw_unchecked,
 * g1.rs:465-470) checks the curve equation only.  Any failing point fails the call
 * with MZK_ERR_INVALID naming the lowest failing index (the reference returns an
 * io::Error or panics).  The file header (k) and the G2 points are the caller's. */
int mzk_srs_read_points(mzk_ctx *ctx, const uint8_t *bytes, size_t count, int format, mzk_srs **out);
/* The point arrays of ParamsKZG::write_custom (params.rs:259-281): bases[offset,
 * offset+count) encoded to host memory (48 B per point for Processed, else 96 B). */
int mzk_srs_write_points(mzk_ctx *ctx, const mzk_srs *srs, size_t offset, size_t count, int format,
                         uint8_t *bytes);
size_t mzk_srs_len(const mzk_srs *srs);
/* Device pointer to the n x 12 u64 affine bases. */
const void *mzk_srs_bases_dev(const mzk_srs *srs);
/* Copies bases[offset, offset+count) to host memory. */
int mzk_srs_download(mzk_srs *srs, size_t offset, size_t count, uint64_t *out);
/* Waits for all work on the SRS's device, then frees the bases (valid also after the
 * creating context is destroyed). */
void mzk_srs_destroy(mzk_srs *srs);
/* KZG commit against resident bases (kzg/mod.rs:82-108 commit_many -> msm_specific):
 * sum_i coeffs[i] * g[offset + i], coeffs device-resident.  The bases are read-only after
 * creation: any context on the same device may commit against them (the work runs on
 * ctx's stream with ctx's scratch), so several contexts can share one SRS. */
int mzk_msm_srs_dev(mzk_ctx *ctx, const mzk_srs *srs, size_t offset, const void *d_coeffs,
                    size_t n, uint64_t out[18]);

/* Many commitments at once (the independent commit_many / par_iter batches of
 * prover.rs:619-623, 758-762 and permutation/prover.rs:241-245): one sort, one bucket
 * accumulation and one reduction over all jobs, so uneven sizes share the device.
 * outs receives 18 u64 per job; each result equals mzk_msm_srs_dev on that job. */
typedef struct {
  const void *d_coeffs; /* n device Montgomery scalars */
  size_t n;
  size_t offset;        /* first SRS base used */
} mzk_msm_job;
int mzk_msm_srs_batch_dev(mzk_ctx *ctx, const mzk_srs *srs, const mzk_msm_job *jobs, size_t njobs,
                          uint64_t *outs);

/* ------------------------------------------------------------------ NTT
 * Transforms of every size the field allows, 2^1 .. 2^32 (S = 32, fq.rs:43-48): one pass up
 * to 2^11, two up to 2^22, three above; device memory permitting.
 *
 * Replaces curves/src/fft.rs:49 `best_fft(a: &mut [G], omega, log_n)` for G = Fq:
 * in-place A_j = sum_i a_i omega^(ij), natural order, host pointer, n = 2^log_n.
 * omega must be a primitive 2^log_n-th root of unity (not checked, as in the reference). */
int mzk_best_fft(mzk_ctx *ctx, uint64_t *a, const uint64_t omega[4], uint32_t log_n);
/* Replaces curves/src/fft.rs:31 `compute_twiddles(omega, log_n) -> Vec<F>`:
 * out[i] = omega^i for i < 2^(log_n - 1) (host memory), 1 <= log_n <= 32. */
int mzk_compute_twiddles(mzk_ctx *ctx, const uint64_t omega[4], uint32_t log_n, uint64_t *out);
/* Replaces curves/src/fft.rs:56 `best_fft_with_twiddles(a, twiddles, log_n)` for G = Fq
 * (as EvaluationDomain::fft / ifft call it, domain.rs:389-396): in place on 2^log_n host
 * elements.  twiddles = compute_twiddles(omega, log_n) (2^(log_n - 1) elements); the
 * transform is the DFT by omega = twiddles[1] (-1 for log_n = 1).  The table is checked
 * where that is cheap (twiddles[0] = 1, omega^(n/2) = -1, the last entry = omega^(n/2 - 1)):
 * a table that fails gives MZK_ERR_INVALID. */
int mzk_best_fft_with_twiddles(mzk_ctx *ctx, uint64_t *a, const uint64_t *twiddles, uint32_t log_n);
/* Replaces curves/src/fft.rs:110 `fft_coeff_to_extended(a, twiddles, log_n, n_real)`: the
 * same DFT of a whose entries from n_real on are zero, as the reference's pruned DIF
 * requires (fft.rs:126-143); only a[0, n_real) is read, all 2^log_n entries are written. */
int mzk_fft_coeff_to_extended(mzk_ctx *ctx, uint64_t *a, const uint64_t *twiddles, uint32_t log_n,
                              size_t n_real);
/* best_fft on device memory with a caller's root: d_out = the DFT by omega of d_in's
 * first n_real elements zero-padded to 2^log_n (n_real = 2^log_n for a plain best_fft;
 * d_in == d_out allowed).  Twiddle tables are kept on the context per (size, root). */
int mzk_best_fft_dev(mzk_ctx *ctx, const void *d_in, void *d_out, const uint64_t omega[4], uint32_t log_n,
                     uint64_t n_real);

/* Replaces proofs/src/poly/domain.rs:34-167 `EvaluationDomain::new(j, k)`:
 * k, extended_k = min{e >= k : 2^e >= 2^k (j-1)} <= 32 (domain.rs:64-70), omega tables
 * resident on device. */
typedef struct mzk_domain mzk_domain;
int mzk_domain_create(mzk_ctx *ctx, uint32_t j, uint32_t k, mzk_domain **out);
/* Waits for all work on the domain's device, then frees its tables (valid also after the
 * creating context is destroyed). */
void mzk_domain_destroy(mzk_domain *d);
int mzk_domain_info(const mzk_domain *d, uint32_t *k, uint32_t *extended_k);
/* domain.rs:249-264 lagrange_to_coeff: in place on 2^k device elements. */
int mzk_lagrange_to_coeff_dev(mzk_domain *d, void *d_a);
/* domain.rs:268-277 coeff_to_lagrange: in place on 2^k device elements. */
int mzk_coeff_to_lagrange_dev(mzk_domain *d, void *d_a);
/* domain.rs:285-305 coeff_to_extended: 2^k coefficients in, 2^extended_k evaluations
 * on the zeta coset out (d_in and d_out must not overlap). */
int mzk_coeff_to_extended_dev(mzk_domain *d, const void *d_in, void *d_out);
/* domain.rs:312-328 extended_to_coeff: in place on 2^extended_k device elements. */
int mzk_extended_to_coeff_dev(mzk_domain *d, void *d_a);
/* domain.rs:361-380 divide_by_vanishing_poly: in place on 2^extended_k elements. */
int mzk_divide_by_vanishing_poly_dev(mzk_domain *d, void *d_a);
/* Stand-alone transforms of size 2^log_n (1 <= log_n <= 32) on device memory, with the
 * domain semantics of domain.rs:
 *   flags 0                          : DFT with omega_n of n_real coefficients zero-padded
 *                                      to 2^log_n (coeff_to_lagrange when n_real = 2^log_n);
 *   MZK_NTT_COSET                    : the same on the zeta coset (coeff_to_extended, :285);
 *   MZK_NTT_INVERSE                  : inverse DFT times 1/2^log_n (lagrange_to_coeff, :249);
 *   MZK_NTT_INVERSE | MZK_NTT_COSET  : extended_to_coeff (:312).
 * d_in holds n_real elements (2^log_n for inverses); d_out holds 2^log_n; in place is
 * allowed.  Twiddle tables are built once per (log_n, direction) and kept on the context. */
enum { MZK_NTT_INVERSE = 1, MZK_NTT_COSET = 2 };
typedef struct {
  const void *d_in;
  void *d_out;
  uint64_t n_real;
  uint32_t log_n;
  uint32_t flags;
} mzk_ntt_job;
int mzk_ntt_dev(mzk_ctx *ctx, const mzk_ntt_job *job);
/* Many transforms of mixed sizes in up to three launches (all first passes, then all
 * second passes, then the third passes of transforms above 2^22): the batched form of the
 * prover's per-column par_iter NTTs (prover.rs:276-280, 652-659).  No job's output may
 * overlap another job's input or output (a job's own d_in == d_out is allowed):
 * MZK_ERR_INVALID names the first such pair. */
int mzk_ntt_batch_dev(mzk_ctx *ctx, const mzk_ntt_job *jobs, size_t njobs);
/* No counterpart in the reference: for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops).  The same transform as mzk_ntt_dev, with
 * the same results, calling no hipMalloc / hipFree / allocating routine: the pass buffer
 * and twiddle tables live in a fixed device array the library holds per device (32 MiB +
 * 192 KiB) until the process ends.  log_n <= MZK_NTT_NOALLOC_MAX_LOG, else MZK_ERR_INVALID.
 * Every call on one device shares that array: calls that overlap on a device race, so
 * order them (one stream).  No graph is captured and no kernel timing is recorded. */
#define MZK_NTT_NOALLOC_MAX_LOG 20
int mzk_ntt_dev_noalloc(mzk_ctx *ctx, const mzk_ntt_job *job);
/* Host-pointer forms of the four transforms. */
int mzk_lagrange_to_coeff(mzk_domain *d, uint64_t *a);
int mzk_coeff_to_extended(mzk_domain *d, const uint64_t *in, uint64_t *out);
int mzk_extended_to_coeff(mzk_domain *d, uint64_t *a);

/* ------------------------------------------------------------------ scalar vectors
 * Device vectors of n Montgomery Fr elements (4 u64 each); in-place use allowed
 * unless stated.  Element-wise a (+,-,*) b: op 0 = add, 1 = sub, 2 = mul
 * (Polynomial Add/Sub/Mul, proofs/src/poly/mod.rs). */
int mzk_fr_vec_op(mzk_ctx *ctx, int op, const void *d_a, const void *d_b, void *d_out, size_t n);
/* With a scalar s: op 0: out = a * s; 1: out = a + s; 2: out = out + a * s
 * (Polynomial Mul<F> / Add<F>, mod.rs; the axpy of poly_inner_product, kzg/mod.rs:125-147). */
int mzk_fr_vec_scalar(mzk_ctx *ctx, int op, const void *d_a, const uint64_t s[4], void *d_out,
                      size_t n);
/* a_i <- a_i^-1 in place, zeros left as zero (ff BatchInvert as used by
 * permutation/prover.rs:148 and domain.rs:141). */
int mzk_fr_batch_invert(mzk_ctx *ctx, void *d_a, size_t n);
/* out_i = a_0 * ... * a_i (inclusive products; the grand product of
 * permutation/prover.rs:190-197 is the exclusive form, out shifted by one).  n <= 2^24.
 * d_out must not alias d_a. */
int mzk_fr_prefix_product(mzk_ctx *ctx, const void *d_a, void *d_out, size_t n);
/* Many independent inclusive product scans in one launch of each phase: the running
 * products of every permutation set (permutation/prover.rs:190-197, one per set and
 * proof).  Job j: d_out = prefix products of d_a (n <= 2^24).  Results equal the per-job
 * mzk_fr_prefix_product.  No output may overlap any job's input or another output. */
typedef struct mzk_prefix_job {
  const void *d_a;
  void *d_out;
  size_t n;
} mzk_prefix_job;
int mzk_fr_prefix_product_batch(mzk_ctx *ctx, const mzk_prefix_job *jobs, size_t njobs);
/* sum_i a_i x^i to host memory (eval_polynomial, utils/arithmetic.rs:59-87). */
int mzk_fr_eval(mzk_ctx *ctx, const void *d_a, size_t n, const uint64_t x[4], uint64_t out[4]);
/* njobs evaluations in one launch and one wait: outs[4j..4j+4) = sum_i polys[j][i] x_j^i,
 * x_j = xs[4j..4j+4) (the batched per-query evaluations of prover.rs:820-843 and
 * kzg/mod.rs:259-260).  Each job may have its own length and point. */
int mzk_fr_eval_batch(mzk_ctx *ctx, const void *const *d_polys, const size_t *ns, const uint64_t *xs,
                      size_t njobs, uint64_t *outs);
/* Quotient of a(X) - a(b) by (X - b): n - 1 coefficients into d_q (kate_division,
 * utils/arithmetic.rs:106-125).  n >= 1; d_q must not alias d_a. */
int mzk_fr_kate_div(mzk_ctx *ctx, const void *d_a, size_t n, const uint64_t b[4], void *d_q);
/* Many independent Kate divisions in one launch of each pass, as the multi-open divides
 * every point set's polynomial in parallel (kzg/mod.rs:226-236, each by kate_division,
 * utils/arithmetic.rs:106-125).  Job j: d_q = quotient of d_a (n coefficients) by
 * (X - b).  Results equal the per-job mzk_fr_kate_div; n = 1 jobs write nothing.  No
 * output may overlap any job's input or another job's output (MZK_ERR_INVALID). */
typedef struct mzk_kate_job {
  const void *d_a;
  size_t n;
  uint64_t b[4];
  void *d_q;
} mzk_kate_job;
int mzk_fr_kate_div_batch(mzk_ctx *ctx, const mzk_kate_job *jobs, size_t njobs);
/* Delta encodings of a Lagrange vector for commits in the delta bases (poly/mod.rs:290-335,
 * Polynomial::to_delta / to_double_delta): order 1, out_0 = a_0, out_i = a_i - a_(i-1);
 * order 2, out_0 = a_0, out_1 = a_1 - 2 a_0, out_i = a_i - 2 a_(i-1) + a_(i-2).  Committed
 * against the suffix-sum bases (mzk_srs_suffix_sum of g_lagrange, once or twice) they give
 * the Lagrange commitment, with zero entries for constant (or linear) runs.  out must not
 * overlap a. */
int mzk_fr_to_delta_dev(mzk_ctx *ctx, const void *d_a, void *d_out, size_t n, int order);
/* out_i = x^i for i < n (powers, kzg/mod.rs:151-167 and the domain's omega tables). */
int mzk_fr_powers(mzk_ctx *ctx, const uint64_t x[4], size_t n, void *d_out);

/* ------------------------------------------------------------------ quotient numerator
 * evaluate_numerator (proofs/src/plonk/evaluation.rs:818-940) fused with
 * divide_by_vanishing_poly (proofs/src/poly/domain.rs:361-380) for a circuit whose
 * only gate is the standard one, q_m a b + q_l a + q_r b - q_o c + q_c, and whose
 * permutation argument covers ncols columns in sets of `chunk`: one pass over the
 * extended-domain (coset) evaluations, element i of
 *   (((gate y + (1 - z_0) l_0) y + (z_l^2 - z_l) l_last) y + sum over sets s > 0 of
 *    (z_s - z_(s-1)(omega^last X)) l_0 ... ) y + (left_s - right_s) l_active ...
 * times t_evals[i mod t_count], in the reference's order of identities and powers of y:
 *   left_s  = z_s(omega X) prod_j (col_j + beta sigma_j + gamma),
 *   right_s = z_s(X) prod_j (col_j + beta delta^j X + gamma),
 * with X = coset_x[i] (zeta omega_ext^i).  Rotations are index shifts on the extended
 * domain.  All vectors are n_ext Montgomery scalars; scalars are Montgomery. */
typedef struct {
  uint64_t n_ext;
  uint32_t rot_next;      /* index shift of omega X: 2^(extended_k - k) */
  uint32_t rot_last;      /* index shift of omega^-(blinding+1) X, as a positive count */
  uint32_t ncols, nsets, chunk, t_count;
  const void *a, *b, *c, *qm, *ql, *qr, *qo, *qc;
  const void *col[8];
  const void *sigma[8];
  const void *z[4];
  const void *l0, *l_last, *l_active, *coset_x;
  uint64_t y[4], beta[4], gamma[4];
  uint64_t delta_beta[8][4];   /* beta delta^j for column j */
  uint64_t t_evals[16][4];     /* 1 / ((zeta omega_ext^i)^n - 1), i < t_count */
} mzk_plonk_numerator;
int mzk_plonk_numerator_dev(mzk_ctx *ctx, const mzk_plonk_numerator *args, void *d_out);
/* The permutation argument's per-set products (proofs/src/plonk/permutation/prover.rs:
 * 124-164), numerator and denominator in one pass, for many sets at once:
 *   den[i] = prod_j (values_j[i] + beta sigma_j[i] + gamma),
 *   num[i] = prod_j (values_j[i] + delta_beta_j omega^i + gamma),
 * with delta_beta_j = beta delta^(set chunk + j) and omega_powers[i] = omega^i, i < n.
 * The reference then inverts den (batch_invert) and multiplies num into it; z is the
 * running product of num / den. */
#define MZK_PERM_MAX_CHUNK 4
typedef struct {
  const void *values[MZK_PERM_MAX_CHUNK];   /* the set's columns (n Lagrange values each) */
  const void *sigma[MZK_PERM_MAX_CHUNK];    /* their permutation polynomials' values */
  uint64_t delta_beta[MZK_PERM_MAX_CHUNK][4];
  uint64_t beta[4], gamma[4];
  uint32_t ncols;                           /* 1 .. MZK_PERM_MAX_CHUNK */
  void *num, *den;                          /* outputs, n each */
} mzk_perm_set;
int mzk_plonk_perm_products_dev(mzk_ctx *ctx, const void *omega_powers, size_t n, const mzk_perm_set *sets,
                                size_t nsets);
/* sum_j s_j p_j (zero-extended to the longest): poly_inner_product (kzg/mod.rs:125-147)
 * in one launch for up to 64 polynomials; out has max_j len_j elements. */
int mzk_fr_lincomb_dev(mzk_ctx *ctx, const void *const *d_polys, const size_t *lens, const uint64_t *scalars,
                       size_t npolys, void *d_out, size_t out_len);

/* ------------------------------------------------------------------ LogUp lookups
 * compute_multiplicities (proofs/src/plonk/logup/prover.rs:411-477): for the first
 * usable_rows rows, m_i = (number of selected input rows, over all ninputs compressed input
 * columns, whose value is table_i) / (number of table rows holding table_i); rows from
 * usable_rows on take blinding[0 .. n - usable_rows) (host memory).  Inputs are counted where
 * the selector is non-zero.  A selected input value missing from the table gives
 * MZK_ERR_INVALID naming its column and row (the reference panics).  All vectors hold n
 * device Montgomery values; d_m must not overlap the others.  Up to 32 input columns. */
int mzk_logup_multiplicities_dev(mzk_ctx *ctx, const void *d_selector, const void *const *d_inputs, size_t ninputs,
                                 const void *d_table, size_t n, size_t usable_rows, const uint64_t *blinding,
                                 void *d_m);
/* compute_logderivative (logup/prover.rs:189-310): with 1/(v + beta) for every input and
 * table value (one batch inversion, zeros kept), helper column k = sum over chunk k's
 * inputs of 1/(f + beta); per row l_i = selector_i sum_k helper_k,i - m_i / (table_i + beta);
 * aggregator Z_0 = 0, Z_j = l_0 + ... + l_(j-1) for j < n - blinding_factors, then the
 * blinding values.  Outputs must not overlap inputs or each other. */
typedef struct {
  size_t n;                        /* rows */
  size_t blinding_factors;         /* Z's last rows take the blinding values */
  const void *selector, *multiplicities, *table;
  const void *const *inputs;       /* compressed input columns, chunk after chunk (<= 32) */
  size_t ninputs;
  const uint32_t *chunk_sizes;     /* input columns per chunk; they sum to ninputs */
  size_t nchunks;
  uint64_t beta[4];
  const uint64_t *blinding;        /* host memory, blinding_factors values */
  void *const *helpers;            /* out: one column of n values per chunk */
  void *aggregator;                /* out: n values */
} mzk_logup_args;
int mzk_logup_logderivative_dev(mzk_ctx *ctx, const mzk_logup_args *args);
/* Inclusive running sums out_i = a_0 + ... + a_i (n <= 2^24; the LogUp aggregator's scan,
 * logup/prover.rs:269-278).  d_out must not alias d_a. */
int mzk_fr_prefix_sum(mzk_ctx *ctx, const void *d_a, void *d_out, size_t n);

/* ------------------------------------------------------------------ native prover
 * create_proof (proofs/src/plonk/prover.rs:432-474) with multi_open
 * (proofs/src/poly/kzg/mod.rs:110-289) for the synthetic circuit (advice a..e, fixed
 * q_m..q_c, gate a b q_m + a q_l + b q_r - c q_o + q_c, all advice columns in the
 * permutation, quotient degree 4, 6 blinding rows), run for `nproofs` proofs of one
 * size in lockstep: each phase is one batched MSM / NTT / evaluation call for the
 * whole group.  The key's vectors are device-resident Montgomery scalars (n = 2^k
 * values, or 2^(k+2) for the coset forms), as the Python keygen builds them.  advice
 * holds nproofs x 5 device pointers to the Lagrange columns (rows n - 7.. are replaced
 * by blinding values drawn from the seeded stream of rng_seeds[p]).  Writes nproofs x
 * MZK_PLONK_PROOF_BYTES.  The columns' transforms to coefficients and to the extended
 * coset run on a second stream of the context (created on first use, released with the
 * context), overlapping the advice and permutation commitments.  phase_ms, if not null,
 * receives 8 cumulative times (ms) at the ends of: advice commitments, permutation
 * commitments, the wait for the columns' transforms, (the same point again), the quotient
 * numerator, the quotient commitments, the evaluations, the multi-open. */
#define MZK_PLONK_PROOF_BYTES 1360
typedef struct {
  uint32_t k;
  const mzk_srs *g, *g_lagrange;                 /* monomial (>= 2^k) and Lagrange (2^k) bases */
  const void *fixed_polys[5], *fixed_cosets[5];
  const void *sigma_values[5], *sigma_polys[5], *sigma_cosets[5];
  const void *l0, *l_last, *l_active, *coset_x;  /* extended-domain vectors */
  const void *omega_powers;                      /* omega^i, i < n */
  uint64_t transcript_repr[4];                   /* VK hash (Montgomery) */
  /* suffix sums of g_lagrange (2^k): when set, the permutation products z are committed as
   * z.to_delta() against them (permutation/prover.rs:236-245) -- the same points, with the
   * constant runs of z skipped; null commits z densely against g_lagrange */
  const mzk_srs *g_lagrange_delta;
} mzk_plonk_key;
int mzk_plonk_prove(mzk_ctx *ctx, const mzk_plonk_key *key, size_t nproofs, const void *const *advice,
                    const uint64_t *rng_seeds, uint8_t *proofs, double *phase_ms);

/* ------------------------------------------------------------------ synthetic inputs
 * Element i of the seeded stream (seed, stream): eight splitmix64 words
 * w_j = splitmix64(key + 8 (offset + i) + j), key = splitmix64(seed ^ stream * 0xD1B54A32D192ED03),
 * reduced as Fq::from_uniform_bytes (fq.rs:712-736).  Writes n Montgomery elements. */
int mzk_fr_from_seed_dev(mzk_ctx *ctx, uint64_t seed, uint64_t stream, uint64_t offset, size_t n,
                         void *d_out);

/* ------------------------------------------------------------------ kernel timing
 * When enabled, every kernel the context launches is bracketed by HIP events on
 * its stream.  mzk_prof_read waits for them and writes one line per kernel name,
 * "<name> <total ms> <launches>\n", then clears the totals.  Enabling or disabling
 * also clears them.  Timing adds event overhead: leave it off for timed runs. */
int mzk_prof_enable(mzk_ctx *ctx, int on);
/* ------------------------------------------------------------------ launch graphs
 * No reference counterpart (the reference's CPU path has no launches).  The fixed
 * device launch sequences of a call -- an MSM's sort, accumulation and reduction
 * kernels, a batch of NTTs' two passes -- run as HIP graphs, one per distinct shape and
 * buffer set: a shape's first call runs eagerly, its second is captured, later ones
 * replay (a prover's rounds repeat the same shapes on the same buffers proof after
 * proof; one-off shapes skip the capture cost and evict nothing).  Up to 48 are held; a
 * held graph is replaced only after 384 calls without use, so a working set of more
 * shapes keeps 48 of them and runs the rest eagerly, with no recapture and no wait.
 * On by default; off replays
 * nothing and frees the graphs.  Results are identical either way; the kernel profiler
 * bypasses graphs while it is on.  out = {graphs captured, replays, graphs held}. */
int mzk_ctx_set_graphs(mzk_ctx *ctx, int on);
int mzk_ctx_graph_stats(mzk_ctx *ctx, uint64_t out[3]);
int mzk_prof_read(mzk_ctx *ctx, char *buf, size_t len);
/* Plan of the last MSM run on the context: out = {window bits c, windows W, bucket
 * sets, references per accumulation thread K, non-zero window digits M, n}. */
int mzk_msm_last_plan(mzk_ctx *ctx, uint64_t out[6]);

/* ------------------------------------------------------------------ autotuner
 * No reference counterpart (the reference fixes c = ceil(ln n), msm.rs:464-470, and
 * its FFT recursion depth by thread count, fft.rs:49-101).  The planners' choices --
 * resident-SRS MSM path (per-window or shared buckets), window bits c, accumulation
 * threads per CU (references per thread), NTT four-step split and workgroup target --
 * come from cost models unless a process-wide table, filled by the tuner or imported,
 * says otherwise.  Every plan returns the same point / DFT values; the tuner checks
 * that for each candidate it times. */
/* Times every MSM candidate (c in {4, 8, 16} x {256, 512, 1024} threads/CU, and the
 * shared path where srs has shifted tables) on seeded scalars against srs at n = 2^lo
 * .. 2^hi (best of reps) and keeps the fastest per size. */
int mzk_tune_msm(mzk_ctx *ctx, const mzk_srs *srs, uint32_t log_lo, uint32_t log_hi, uint32_t reps);
/* The same for forward NTTs of 2^lo .. 2^hi (sizes above 2^11: splits within two of
 * the even split x workgroup targets {256, 1024, 4096}). */
int mzk_tune_ntt(mzk_ctx *ctx, uint32_t log_lo, uint32_t log_hi, uint32_t reps);
/* The table as text lines "msm <log_n> <path> <c> <tpc_window> <tpc_shared>" and
 * "ntt <log_n> <l1> <groups>" (zero / -1 = model); *needed = bytes incl. the NUL.
 * len = 0 only reports the size. */
int mzk_tune_export(char *buf, size_t len, size_t *needed);
/* Merges such text into the table ('#' starts a comment line); all or nothing. */
int mzk_tune_import(const char *text);
/* Back to the models. */
void mzk_tune_reset(void);

/* ------------------------------------------------------------------ diagnostics
 * Element-wise field / curve operations for parity tests (device pointers, n elements):
 *   field 0 = fr (8 u32 / element), 1 = fp (12 u32 / element);
 *   op 0 = mul (the product-scanning form every kernel uses), 1 = add, 2 = sub,
 *   3 = square, 4 = from_mont, 5 = mul by the operand-scanning (CIOS) form. */
int mzk_diag_field_op(mzk_ctx *ctx, int field, int op, const void *d_a, const void *d_b, void *d_out,
                      size_t n);
/* G1: op 0 = affine + affine (through XYZZ), 1 = 2 * affine, 2 = XYZZ + XYZZ of the
 * two affine inputs each doubled first; outputs Jacobian (18 u64 / element). */
int mzk_diag_g1_op(mzk_ctx *ctx, int op, const void *d_a, const void *d_b, void *d_out, size_t n);
/* Micro-benchmarks (time in ms from HIP events):
 *   which 0: independent 32x32+64 multiply-adds, result = mads issued;
 *   which 1: fp CIOS products, many independent chains (throughput);
 *   which 2: fp CIOS products, one wave, one dependent chain (latency);
 *   which 3: fr CIOS products, many chains (throughput);
 *   which 4: dependent 32x32+64 multiply-add chains with no other instruction
 *            (v_mad_u64_u32 issue rate: the INT32 multiply roof of the chip);
 *   which 5: fp product-scanning (default) products, many chains (throughput);
 *   which 6: fr product-scanning (default) products, many chains (throughput);
 *   which 7: fp products in the radix-2^28 form of the MSM accumulation, many chains;
 *   which 8: XYZZ mixed additions, 32-bit form (one accumulator per thread);
 *   which 9: XYZZ mixed additions, radix-2^28 form;
 *   which 10: dependent 32-bit multiply-lows (v_mul_lo_u32 issue rate). */
int mzk_diag_bench(mzk_ctx *ctx, int which, int blocks, int threads, int iters, double *ms,
                   double *ops);

#ifdef __cplusplus
}
#endif
#endif /* MZK_GPU_H */
