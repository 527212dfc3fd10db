/* Drives every entry point of the oracle's C restatement (oracle/cpu_ref.c, compiled into
 * this program) on seeded inputs, so that a sanitizer build (AddressSanitizer +
 * UndefinedBehaviorSanitizer) sees each code path: MSM below and above the batch-affine
 * threshold (8104 terms) with zero, repeated and negated bases, every FFT form, the
 * domain transforms and the vector routines.  Results are cross-checked against each
 * other (msm_best = msm_parallel = msm_specific = [sum s_i tau^i] G; inverse transforms
 * undo forward ones; a batch-inverted vector times the original is one); prints "ok"
 * and exits 0, or names the first mismatch and exits 1.  Test infrastructure only. */
#include "../../oracle/cpu_ref.c"

#include <stdio.h>

static int fails = 0;
#define CHECK(cond, what)                          \
  do {                                             \
    if (!(cond)) {                                 \
      fprintf(stderr, "mismatch: %s\n", what);     \
      fails++;                                     \
    }                                              \
  } while (0)

static int aff_eq(const u64 *jac_a, const u64 *jac_b) {
  u64 a[12], b[12];
  cpu_ref_g1_to_affine(jac_a, a);
  cpu_ref_g1_to_affine(jac_b, b);
  return !memcmp(a, b, sizeof a);
}

static void msm_checks(size_t n, const u64 *tau) {
  u64 *bases = (u64 *)calloc(12 * (n + 1), 8);
  u64 *sc = (u64 *)calloc(4 * (n + 1), 8);
  cpu_ref_srs_bases(tau, n, bases);
  cpu_ref_fr_from_seed(77, n, 0, n, sc);
  if (n > 10) {
    memset(sc + 4 * 3, 0, 32);                       /* a zero scalar */
    memcpy(bases + 12 * 5, bases + 12 * 4, 96);      /* a repeated base: doubling in a bucket */
    memcpy(sc + 4 * 5, sc + 4 * 4, 32);
  }
  u64 best[18], par[18], spec[18], ser[18] = {0}, expect[18];
  cpu_ref_msm_best(sc, bases, n, best);
  cpu_ref_msm_parallel(sc, bases, n, par);
  cpu_ref_msm_specific(sc, bases, n, spec);
  cpu_ref_msm_serial(sc, bases, n, ser);
  CHECK(aff_eq(best, par), "msm_best vs msm_parallel");
  CHECK(aff_eq(best, spec), "msm_best vs msm_specific");
  CHECK(aff_eq(best, ser), "msm_best vs msm_serial from the identity");
  if (n <= 10) {   /* the closed form holds for untouched SRS bases */
    u64 s[4];
    cpu_ref_poly_eval(sc, n, tau, s);
    cpu_ref_g1_mul_generator(s, expect);   /* Jacobian */
    CHECK(aff_eq(best, expect), "msm_best vs the closed form");
  }
  free(bases);
  free(sc);
}

static void fft_checks(unsigned log_n) {
  size_t n = (size_t)1 << log_n;
  u64 *a = (u64 *)malloc(32 * n), *b = (u64 *)malloc(32 * n), *c = (u64 *)malloc(32 * n);
  u64 *tw = (u64 *)malloc(32 * (n / 2 + 1)), w[4];
  cpu_ref_fr_from_seed(5, log_n, 0, n, a);
  memcpy(b, a, 32 * n);
  cpu_ref_ntt(b, log_n);
  memcpy(c, a, 32 * n);
  cpu_ref_root_of_unity(log_n, w);
  cpu_ref_compute_twiddles(w, log_n, tw);
  cpu_ref_best_fft_with_twiddles(c, tw, log_n);
  CHECK(!memcmp(b, c, 32 * n), "ntt vs best_fft_with_twiddles");
  memcpy(c, a, 32 * n);
  memset(c + 4 * (n / 2), 0, 32 * (n / 2));
  u64 *d = (u64 *)malloc(32 * n);
  memcpy(d, c, 32 * n);
  cpu_ref_fft_coeff_to_extended(c, tw, log_n, n / 2);
  cpu_ref_ntt(d, log_n);
  CHECK(!memcmp(c, d, 32 * n), "fft_coeff_to_extended vs ntt of the padded input");
  cpu_ref_intt(b, log_n);
  CHECK(!memcmp(a, b, 32 * n), "intt(ntt(a))");
  /* coset extension by 4 and back */
  u64 *ext = (u64 *)malloc(32 * 4 * n);
  cpu_ref_coeff_to_extended(a, log_n, log_n + 2, ext);
  u64 *ext2 = (u64 *)malloc(32 * 4 * n);
  memcpy(ext2, ext, 32 * 4 * n);
  cpu_ref_divide_by_vanishing_poly(ext2, log_n, log_n + 2);
  cpu_ref_extended_to_coeff(ext, log_n + 2);
  CHECK(!memcmp(ext, a, 32 * n), "extended_to_coeff(coeff_to_extended(a))");
  free(a); free(b); free(c); free(d); free(tw); free(ext); free(ext2);
}

static void vector_checks(size_t n) {
  u64 *a = (u64 *)malloc(32 * n), *b = (u64 *)malloc(32 * n), *o = (u64 *)malloc(32 * n);
  u64 *inv = (u64 *)malloc(32 * n), *q = (u64 *)malloc(32 * n), *pw = (u64 *)malloc(32 * n);
  u64 s[4], one[4], e[4];
  cpu_ref_fr_from_seed(9, 1, 0, n, a);
  cpu_ref_fr_from_seed(9, 2, 0, n, b);
  memset(a + 4 * 2, 0, 32);                        /* a zero stays zero under inversion */
  cpu_ref_fr_from_seed(9, 3, 0, 1, s);
  for (int op = 0; op < 3; op++) cpu_ref_fr_vec_op(op, a, b, o, n);
  for (int op = 0; op < 3; op++) cpu_ref_fr_vec_scalar(op, a, s, o, n);
  memcpy(inv, a, 32 * n);
  cpu_ref_fr_batch_invert(inv, n);
  cpu_ref_fr_vec_op(2, a, inv, o, n);
  cpu_ref_fr_powers(s, 1, one);                    /* s^0 = 1 */
  for (size_t i = 0; i < n; i++)
    CHECK(i == 2 ? !(o[4 * i] | o[4 * i + 1] | o[4 * i + 2] | o[4 * i + 3]) : !memcmp(o + 4 * i, one, 32),
          "a * batch_invert(a)");
  cpu_ref_fr_prefix_product(b, o, n);
  cpu_ref_fr_kate_div(a, n, s, q);
  cpu_ref_fr_powers(s, n, pw);
  cpu_ref_poly_eval(a, n, s, e);
  u64 *aff = (u64 *)malloc(96 * 8);
  cpu_ref_g1_mul_generator_many(b, 8, aff);
  free(a); free(b); free(o); free(inv); free(q); free(pw); free(aff);
}

static void logup_checks(void) {
  enum { N = 64, BF = 5, U = N - BF - 1, NC = 3 };
  u64 *table = (u64 *)malloc(32 * N), *sel = (u64 *)malloc(32 * N), *m = (u64 *)malloc(32 * N);
  u64 *in[NC], *h[2], *z = (u64 *)malloc(32 * N), bl[(N - U) * 4], beta[4];   /* m takes N - U */
  cpu_ref_fr_from_seed(31, 1, 0, N, table);
  for (int i = U / 4; i < U; i++) memcpy(table + 4 * i, table + 4 * (i % (U / 4)), 32);   /* duplicates */
  for (int i = 0; i < N; i++) { memset(sel + 4 * i, 0, 32); if (i % 3) cpu_ref_fr_powers(beta, 1, sel + 4 * i); }
  for (int c = 0; c < NC; c++) {
    in[c] = (u64 *)malloc(32 * N);
    for (int i = 0; i < N; i++) memcpy(in[c] + 4 * i, table + 4 * ((i * 7 + c) % U), 32);
  }
  cpu_ref_fr_from_seed(31, 2, 0, N - U, bl);
  cpu_ref_fr_from_seed(31, 3, 0, 1, beta);
  CHECK(cpu_ref_logup_multiplicities(sel, (const u64 *const *)in, NC, table, N, U, bl, m) == 0, "logup multiplicities");
  uint32_t cs[2] = {2, 1};
  h[0] = (u64 *)malloc(32 * N);
  h[1] = (u64 *)malloc(32 * N);
  cpu_ref_logup_logderivative(N, BF, sel, m, table, (const u64 *const *)in, NC, cs, 2, beta, bl, h, z);
  CHECK(!(z[4 * U] | z[4 * U + 1] | z[4 * U + 2] | z[4 * U + 3]), "logup balance");
  memcpy(in[1] + 4 * 2, bl, 32);   /* a value outside the table on a selected row */
  CHECK(cpu_ref_logup_multiplicities(sel, (const u64 *const *)in, NC, table, N, U, bl, m) == 1 + N + 2,
        "logup missing value");
  for (int c = 0; c < NC; c++) free(in[c]);
  free(table); free(sel); free(m); free(z); free(h[0]); free(h[1]);
}

int main(void) {
  cpu_ref_set_threads(4);
  u64 tau[4];
  cpu_ref_fr_from_seed(1, 0, 0, 1, tau);
  size_t sizes[] = {0, 1, 3, 10, 40, 9000};
  for (size_t i = 0; i < sizeof sizes / sizeof sizes[0]; i++) msm_checks(sizes[i], tau);
  for (unsigned ln = 1; ln <= 12; ln += 1) fft_checks(ln);
  vector_checks(1000);
  logup_checks();
  u64 g[12];
  cpu_ref_g1_generator(g);
  uint8_t rep[32] = {0xff, 0x01, 0x80};
  for (int c = 1; c <= 16; c++) (void)cpu_ref_booth_index(0, c, rep);
  (void)cpu_ref_window_size(12345);
  if (fails) return 1;
  puts("ok");
  return 0;
}
