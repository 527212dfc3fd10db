/*
 * Oracle: C restatement of the reference CPU path for the prover's MSM / NTT.
 *
 * TEST INFRASTRUCTURE ONLY.  Built into oracle/libcpu_ref.so and loaded (via
 * oracle/cpu_ref.py) only by tests/, by the smoke check and by the
 * cpu_baseline leg of bench.py, as the checker and as the timed CPU baseline.
 * The product library never links or calls it.
 *
 * Restated reference routines (file:line into the reference tree):
 *   fr / fp Montgomery arithmetic ........ curves/src/bls12_381/fq.rs:309-343, fp.rs:430-505
 *                                          (arithmetic itself lives in blst 0.3.15, absent;
 *                                          written here as CIOS Montgomery, HAC 14.36)
 *   Fq::from_uniform_bytes ............... fq.rs:712-736
 *   G1 Jacobian add / mixed add / double . g1.rs:165-240, 582-590, 693 (blst; textbook
 *                                          formulas add-2007-bl / madd-2007-bl / dbl-2009-l)
 *   get_booth_index ...................... curves/src/msm.rs:13-55
 *   batch_add + Schedule ................. msm.rs:58-144, 273-333 (BATCH_SIZE = 64)
 *   msm_serial / msm_parallel / msm_best . msm.rs:338-532
 *   msm_specific (zero filter) ........... proofs/src/poly/kzg/msm.rs:174-201
 *   compute_twiddles / best_fft .......... curves/src/fft.rs:31-101, 202-236
 *   fft_coeff_to_extended (pruned DIF) ... fft.rs:110-198
 *   EvaluationDomain transforms .......... proofs/src/poly/domain.rs:54-414
 *   unsafe_setup bases g_i = [tau^i]G .... proofs/src/poly/kzg/params.rs:154-172
 *
 * Pinned by tests/test_oracle_c.py against the Python oracle (itself pinned by
 * the reference's KATs) and against brute force / closed forms.
 *
 * Parallelism: pthreads stand in for the rayon pool (msm_best runs its windows
 * in parallel, msm_parallel splits points into num_threads chunks, the FFT
 * recursion forks at the top levels).
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef uint64_t u64;
typedef unsigned __int128 u128;

/* ------------------------------------------------------------------ fr */
static const u64 FR_P[4] = {0xffffffff00000001ull, 0x53bda402fffe5bfeull,
                            0x3339d80809a1d805ull, 0x73eda753299d7d48ull};
static const u64 FR_INV = 0xfffffffeffffffffull;
static const u64 FR_ONE[4] = {0x00000001fffffffeull, 0x5884b7fa00034802ull,
                              0x998c4fefecbc4ff5ull, 0x1824b159acc5056full};
static const u64 FR_R2[4] = {0xc999e990f3f29c6dull, 0x2b6cedcb87925c23ull,
                             0x05d314967254398full, 0x0748d9d99f59ff11ull};
static const u64 FR_R3[4] = {0xc62c1807439b73afull, 0x1b3e0d188cf06990ull,
                             0x73d13c71c7b5f418ull, 0x6e2a5bb9c8db33e9ull};
static const u64 FR_ZETA[4] = {0x92d9090b093011d2ull, 0xfc9cbd719d6aa073ull,
                               0xc1f14ef0cd65a1a6ull, 0x017f6d35e72fcdebull};
static const u64 FR_ROOT[4] = {0xb9b58d8c5f0e466aull, 0x5b1b4c801819d7ecull,
                               0x0af53ae352a31e64ull, 0x5bf3adda19e9b27bull};

static int geq_n(const u64 *a, const u64 *p, int n) {
  for (int i = n - 1; i >= 0; i--) {
    if (a[i] > p[i]) return 1;
    if (a[i] < p[i]) return 0;
  }
  return 1;
}
static u64 sub_n(u64 *r, const u64 *a, const u64 *b, int n) {
  u64 borrow = 0;
  for (int i = 0; i < n; i++) {
    u128 d = (u128)a[i] - b[i] - borrow;
    r[i] = (u64)d;
    borrow = (u64)(d >> 64) & 1;
  }
  return borrow;
}
static u64 add_n(u64 *r, const u64 *a, const u64 *b, int n) {
  u64 carry = 0;
  for (int i = 0; i < n; i++) {
    u128 s = (u128)a[i] + b[i] + carry;
    r[i] = (u64)s;
    carry = (u64)(s >> 64);
  }
  return carry;
}

/* CIOS Montgomery product for an n-limb modulus; inputs < 2^(64n) with a*b < p*R
 * give a fully reduced output. */
static inline void mont_mul_n(u64 *r, const u64 *a, const u64 *b, const u64 *p, u64 inv, int n) {
  u64 t[8] = {0};
  for (int i = 0; i < n; i++) {
    u128 c = 0;
    for (int j = 0; j < n; j++) {
      c = (u128)a[j] * b[i] + t[j] + (u64)(c >> 64);
      t[j] = (u64)c;
    }
    u128 s = (u128)t[n] + (u64)(c >> 64);
    t[n] = (u64)s;
    t[n + 1] = (u64)(s >> 64);
    u64 m = t[0] * inv;
    c = (u128)m * p[0] + t[0];
    for (int j = 1; j < n; j++) {
      c = (u128)m * p[j] + t[j] + (u64)(c >> 64);
      t[j - 1] = (u64)c;
    }
    s = (u128)t[n] + (u64)(c >> 64);
    t[n - 1] = (u64)s;
    t[n] = t[n + 1] + (u64)(s >> 64);
  }
  if (t[n] || geq_n(t, p, n)) sub_n(t, t, p, n);
  memcpy(r, t, 8 * n);
}

typedef struct { u64 l[4]; } fr_t;
static inline void fr_mul(fr_t *r, const fr_t *a, const fr_t *b) { mont_mul_n(r->l, a->l, b->l, FR_P, FR_INV, 4); }
static inline void fr_add(fr_t *r, const fr_t *a, const fr_t *b) {
  u64 c = add_n(r->l, a->l, b->l, 4);
  if (c || geq_n(r->l, FR_P, 4)) sub_n(r->l, r->l, FR_P, 4);
}
static inline void fr_sub(fr_t *r, const fr_t *a, const fr_t *b) {
  if (sub_n(r->l, a->l, b->l, 4)) add_n(r->l, r->l, FR_P, 4);
}
static inline int fr_is_zero(const fr_t *a) { return !(a->l[0] | a->l[1] | a->l[2] | a->l[3]); }
static void fr_from_mont(fr_t *r, const fr_t *a) {
  fr_t one = {{1, 0, 0, 0}};
  fr_mul(r, a, &one);
}
static void fr_pow_u64(fr_t *r, const fr_t *a, u64 e) {
  fr_t acc, base = *a;
  memcpy(acc.l, FR_ONE, 32);
  while (e) {
    if (e & 1) fr_mul(&acc, &acc, &base);
    fr_mul(&base, &base, &base);
    e >>= 1;
  }
  *r = acc;
}
static void fr_pow_limbs(fr_t *r, const fr_t *a, const u64 *e, int nlimbs) {
  fr_t acc;
  memcpy(acc.l, FR_ONE, 32);
  for (int i = nlimbs * 64 - 1; i >= 0; i--) {
    fr_mul(&acc, &acc, &acc);
    if ((e[i / 64] >> (i % 64)) & 1) fr_mul(&acc, &acc, a);
  }
  *r = acc;
}
static void fr_inv(fr_t *r, const fr_t *a) {
  u64 e[4];
  memcpy(e, FR_P, 32);
  e[0] -= 2;
  fr_pow_limbs(r, a, e, 4);
}
static void fr_from_u64(fr_t *r, u64 v) {
  fr_t t = {{v, 0, 0, 0}};
  fr_t r2;
  memcpy(r2.l, FR_R2, 32);
  fr_mul(r, &t, &r2);
}

/* ------------------------------------------------------------------ fp */
static const u64 FP_P[6] = {0xb9feffffffffaaabull, 0x1eabfffeb153ffffull, 0x6730d2a0f6b0f624ull,
                            0x64774b84f38512bfull, 0x4b1ba7b6434bacd7ull, 0x1a0111ea397fe69aull};
static const u64 FP_INV = 0x89f3fffcfffcfffdull;
static const u64 FP_ONE[6] = {0x760900000002fffdull, 0xebf4000bc40c0002ull, 0x5f48985753c758baull,
                              0x77ce585370525745ull, 0x5c071a97a256ec6dull, 0x15f65ec3fa80e493ull};

typedef struct { u64 l[6]; } fp_t;
static inline void fp_mul(fp_t *r, const fp_t *a, const fp_t *b) { mont_mul_n(r->l, a->l, b->l, FP_P, FP_INV, 6); }
static inline void fp_sqr(fp_t *r, const fp_t *a) { fp_mul(r, a, a); }
static inline void fp_add(fp_t *r, const fp_t *a, const fp_t *b) {
  u64 c = add_n(r->l, a->l, b->l, 6);
  if (c || geq_n(r->l, FP_P, 6)) sub_n(r->l, r->l, FP_P, 6);
}
static inline void fp_sub(fp_t *r, const fp_t *a, const fp_t *b) {
  if (sub_n(r->l, a->l, b->l, 6)) add_n(r->l, r->l, FP_P, 6);
}
static inline void fp_dbl(fp_t *r, const fp_t *a) { fp_add(r, a, a); }
static inline int fp_is_zero(const fp_t *a) {
  return !(a->l[0] | a->l[1] | a->l[2] | a->l[3] | a->l[4] | a->l[5]);
}
static inline int fp_eq(const fp_t *a, const fp_t *b) { return !memcmp(a->l, b->l, 48); }
static inline void fp_neg(fp_t *r, const fp_t *a) {
  if (fp_is_zero(a)) { *r = *a; return; }
  sub_n(r->l, FP_P, a->l, 6);
}
/* Fermat inversion with a 4-bit fixed window. */
static void fp_inv(fp_t *r, const fp_t *a) {
  u64 e[6];
  memcpy(e, FP_P, 48);
  e[0] -= 2;
  fp_t tab[16];
  memcpy(tab[0].l, FP_ONE, 48);
  tab[1] = *a;
  for (int i = 2; i < 16; i++) fp_mul(&tab[i], &tab[i - 1], a);
  fp_t acc;
  memcpy(acc.l, FP_ONE, 48);
  for (int i = 95; i >= 0; i--) {
    for (int k = 0; k < 4; k++) fp_sqr(&acc, &acc);
    int d = (int)((e[i / 16] >> ((i % 16) * 4)) & 15);
    if (d) fp_mul(&acc, &acc, &tab[d]);
  }
  *r = acc;
}

/* ------------------------------------------------------------------ G1 */
typedef struct { fp_t x, y; } g1a_t;          /* affine; identity = all zero (g1.rs:32-34) */
typedef struct { fp_t x, y, z; } g1j_t;       /* Jacobian; identity z = 0 */

static inline int g1a_is_inf(const g1a_t *p) { return fp_is_zero(&p->x) && fp_is_zero(&p->y); }
static inline int g1j_is_inf(const g1j_t *p) { return fp_is_zero(&p->z); }
static inline void g1j_set_inf(g1j_t *p) { memset(p, 0, sizeof *p); }

static void g1j_dbl(g1j_t *r, const g1j_t *p) {       /* dbl-2009-l, a = 0 */
  if (g1j_is_inf(p)) { *r = *p; return; }
  fp_t A, B, C, D, E, F, t;
  fp_sqr(&A, &p->x);
  fp_sqr(&B, &p->y);
  fp_sqr(&C, &B);
  fp_add(&t, &p->x, &B);
  fp_sqr(&t, &t);
  fp_sub(&t, &t, &A);
  fp_sub(&t, &t, &C);
  fp_dbl(&D, &t);
  fp_dbl(&E, &A);
  fp_add(&E, &E, &A);
  fp_sqr(&F, &E);
  fp_t Z3;
  fp_mul(&Z3, &p->y, &p->z);
  fp_dbl(&Z3, &Z3);
  fp_t X3;
  fp_sub(&X3, &F, &D);
  fp_sub(&X3, &X3, &D);
  fp_t Y3;
  fp_sub(&t, &D, &X3);
  fp_mul(&Y3, &E, &t);
  fp_dbl(&C, &C);
  fp_dbl(&C, &C);
  fp_dbl(&C, &C);
  fp_sub(&Y3, &Y3, &C);
  r->x = X3; r->y = Y3; r->z = Z3;
}

static void g1j_add(g1j_t *r, const g1j_t *p, const g1j_t *q) {   /* add-2007-bl */
  if (g1j_is_inf(p)) { *r = *q; return; }
  if (g1j_is_inf(q)) { *r = *p; return; }
  fp_t Z1Z1, Z2Z2, U1, U2, S1, S2, t;
  fp_sqr(&Z1Z1, &p->z);
  fp_sqr(&Z2Z2, &q->z);
  fp_mul(&U1, &p->x, &Z2Z2);
  fp_mul(&U2, &q->x, &Z1Z1);
  fp_mul(&t, &q->z, &Z2Z2);
  fp_mul(&S1, &p->y, &t);
  fp_mul(&t, &p->z, &Z1Z1);
  fp_mul(&S2, &q->y, &t);
  if (fp_eq(&U1, &U2)) {
    if (fp_eq(&S1, &S2)) { g1j_dbl(r, p); return; }
    g1j_set_inf(r);
    return;
  }
  fp_t H, I, J, rr, V;
  fp_sub(&H, &U2, &U1);
  fp_dbl(&I, &H);
  fp_sqr(&I, &I);
  fp_mul(&J, &H, &I);
  fp_sub(&rr, &S2, &S1);
  fp_dbl(&rr, &rr);
  fp_mul(&V, &U1, &I);
  fp_t X3, Y3, Z3;
  fp_sqr(&X3, &rr);
  fp_sub(&X3, &X3, &J);
  fp_sub(&X3, &X3, &V);
  fp_sub(&X3, &X3, &V);
  fp_sub(&t, &V, &X3);
  fp_mul(&Y3, &rr, &t);
  fp_mul(&t, &S1, &J);
  fp_dbl(&t, &t);
  fp_sub(&Y3, &Y3, &t);
  fp_add(&Z3, &p->z, &q->z);
  fp_sqr(&Z3, &Z3);
  fp_sub(&Z3, &Z3, &Z1Z1);
  fp_sub(&Z3, &Z3, &Z2Z2);
  fp_mul(&Z3, &Z3, &H);
  r->x = X3; r->y = Y3; r->z = Z3;
}

static void g1j_from_affine(g1j_t *r, const g1a_t *a) {
  if (g1a_is_inf(a)) { g1j_set_inf(r); return; }
  r->x = a->x; r->y = a->y;
  memcpy(r->z.l, FP_ONE, 48);
}

static void g1j_add_affine(g1j_t *r, const g1j_t *p, const g1a_t *q) {  /* madd-2007-bl */
  if (g1a_is_inf(q)) { *r = *p; return; }
  if (g1j_is_inf(p)) { g1j_from_affine(r, q); return; }
  fp_t Z1Z1, U2, S2, t;
  fp_sqr(&Z1Z1, &p->z);
  fp_mul(&U2, &q->x, &Z1Z1);
  fp_mul(&t, &p->z, &Z1Z1);
  fp_mul(&S2, &q->y, &t);
  if (fp_eq(&U2, &p->x)) {
    if (fp_eq(&S2, &p->y)) { g1j_dbl(r, p); return; }
    g1j_set_inf(r);
    return;
  }
  fp_t H, HH, I, J, rr, V;
  fp_sub(&H, &U2, &p->x);
  fp_sqr(&HH, &H);
  fp_dbl(&I, &HH);
  fp_dbl(&I, &I);
  fp_mul(&J, &H, &I);
  fp_sub(&rr, &S2, &p->y);
  fp_dbl(&rr, &rr);
  fp_mul(&V, &p->x, &I);
  fp_t X3, Y3, Z3;
  fp_sqr(&X3, &rr);
  fp_sub(&X3, &X3, &J);
  fp_sub(&X3, &X3, &V);
  fp_sub(&X3, &X3, &V);
  fp_sub(&t, &V, &X3);
  fp_mul(&Y3, &rr, &t);
  fp_mul(&t, &p->y, &J);
  fp_dbl(&t, &t);
  fp_sub(&Y3, &Y3, &t);
  fp_add(&Z3, &p->z, &H);
  fp_sqr(&Z3, &Z3);
  fp_sub(&Z3, &Z3, &Z1Z1);
  fp_sub(&Z3, &Z3, &HH);
  r->x = X3; r->y = Y3; r->z = Z3;
}

static void g1a_neg(g1a_t *r, const g1a_t *a) { r->x = a->x; fp_neg(&r->y, &a->y); }

static void g1j_to_affine(g1a_t *r, const g1j_t *p) {
  if (g1j_is_inf(p)) { memset(r, 0, sizeof *r); return; }
  fp_t zi, zi2, zi3;
  fp_inv(&zi, &p->z);
  fp_sqr(&zi2, &zi);
  fp_mul(&zi3, &zi2, &zi);
  fp_mul(&r->x, &p->x, &zi2);
  fp_mul(&r->y, &p->y, &zi3);
}

/* ------------------------------------------------------------------ threads */
typedef struct {
  void (*fn)(void *, int);
  void *arg;
  int n, next;
  pthread_mutex_t mu;
} pool_t;
static void *pool_worker(void *v) {
  pool_t *p = (pool_t *)v;
  for (;;) {
    pthread_mutex_lock(&p->mu);
    int i = p->next++;
    pthread_mutex_unlock(&p->mu);
    if (i >= p->n) break;
    p->fn(p->arg, i);
  }
  return NULL;
}
/* Runs fn(arg, i) for i in [0, n) on up to `threads` OS threads. */
static void parallel_for(int n, int threads, void (*fn)(void *, int), void *arg) {
  if (threads <= 1 || n <= 1) { for (int i = 0; i < n; i++) fn(arg, i); return; }
  pool_t pool;
  pool.fn = fn; pool.arg = arg; pool.n = n; pool.next = 0;
  pthread_mutex_init(&pool.mu, NULL);
  int nt = threads < n ? threads : n;
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nt);
  for (int t = 0; t < nt; t++) pthread_create(&th[t], NULL, pool_worker, &pool);
  for (int t = 0; t < nt; t++) pthread_join(th[t], NULL);
  free(th);
  pthread_mutex_destroy(&pool.mu);
}

static int g_threads = 1;
void cpu_ref_set_threads(int t) { g_threads = t < 1 ? 1 : t; }
int cpu_ref_get_threads(void) { return g_threads; }

/* ------------------------------------------------------------------ Booth (msm.rs:13-55) */
static int32_t get_booth_index(int window_index, int window_size, const uint8_t *el) {
  int skip_bits = window_index * window_size - 1;
  if (skip_bits < 0) skip_bits = 0;
  int skip_bytes = skip_bits / 8;
  uint8_t v[4] = {0, 0, 0, 0};
  for (int i = 0; i < 4 && skip_bytes + i < 32; i++) v[i] = el[skip_bytes + i];
  uint32_t tmp = (uint32_t)v[0] | ((uint32_t)v[1] << 8) | ((uint32_t)v[2] << 16) | ((uint32_t)v[3] << 24);
  if (window_index == 0) tmp <<= 1;
  tmp >>= (skip_bits - skip_bytes * 8);
  tmp &= (1u << (window_size + 1)) - 1;
  int sign = (tmp & (1u << window_size)) == 0;
  tmp = (tmp + 1) >> 1;
  if (sign) return (int32_t)tmp;
  return -(int32_t)((~(tmp - 1)) & ((1u << window_size) - 1));
}
int32_t cpu_ref_booth_index(int w, int c, const uint8_t *el) { return get_booth_index(w, c, el); }

/* ------------------------------------------------------------------ batch-affine buckets (msm.rs:58-333) */
#define BATCH_SIZE 64
typedef struct { int inf; g1a_t p; } buck_aff_t;
typedef struct { int kind; g1j_t p; } buck_jac_t;        /* kind 0 = None */
typedef struct { size_t base_idx, buck_idx; int sign; } sched_pt_t;

static void batch_add(int size, buck_aff_t *buckets, const sched_pt_t *pts, const g1a_t *bases) {
  fp_t t[BATCH_SIZE], z[BATCH_SIZE], acc;
  memcpy(acc.l, FP_ONE, 48);
  for (int i = 0; i < size; i++) {
    buck_aff_t *b = &buckets[pts[i].buck_idx];
    const g1a_t *q = &bases[pts[i].base_idx];
    if (b->inf) continue;
    if (fp_eq(&b->p.x, &q->x)) {
      if (fp_eq(&b->p.y, &q->y) ^ !pts[i].sign) {
        fp_t xs, m;
        fp_sqr(&xs, &q->x);
        fp_add(&m, &xs, &xs);
        fp_add(&m, &m, &xs);
        fp_add(&z[i], &b->p.y, &b->p.y);
        fp_mul(&t[i], &acc, &m);
        fp_mul(&acc, &acc, &z[i]);
        continue;
      }
      b->inf = 1;
      continue;
    }
    fp_sub(&z[i], &b->p.x, &q->x);
    fp_t dy;
    if (pts[i].sign) fp_sub(&dy, &b->p.y, &q->y); else fp_add(&dy, &b->p.y, &q->y);
    fp_mul(&t[i], &acc, &dy);
    fp_mul(&acc, &acc, &z[i]);
  }
  fp_inv(&acc, &acc);
  for (int i = size - 1; i >= 0; i--) {
    buck_aff_t *b = &buckets[pts[i].buck_idx];
    const g1a_t *q = &bases[pts[i].base_idx];
    if (b->inf) continue;
    fp_t lambda, x, tmp;
    fp_mul(&lambda, &acc, &t[i]);
    fp_mul(&acc, &acc, &z[i]);
    fp_sqr(&x, &lambda);
    fp_add(&tmp, &b->p.x, &q->x);
    fp_sub(&x, &x, &tmp);
    fp_sub(&tmp, &q->x, &x);
    fp_mul(&tmp, &lambda, &tmp);
    if (pts[i].sign) fp_sub(&b->p.y, &tmp, &q->y); else fp_add(&b->p.y, &tmp, &q->y);
    b->p.x = x;
  }
}

typedef struct {
  buck_aff_t *buckets;
  sched_pt_t set[BATCH_SIZE];
  int ptr;
} schedule_t;

static int sched_contains(const schedule_t *s, size_t buck_idx) {
  for (int i = 0; i < s->ptr; i++) if (s->set[i].buck_idx == buck_idx) return 1;
  return 0;
}
static void sched_execute(schedule_t *s, const g1a_t *bases) {
  if (s->ptr) { batch_add(s->ptr, s->buckets, s->set, bases); s->ptr = 0; }
}
static void sched_add(schedule_t *s, const g1a_t *bases, size_t base_idx, size_t buck_idx, int sign) {
  buck_aff_t *b = &s->buckets[buck_idx];
  if (b->inf) {
    b->inf = 0;
    if (sign) b->p = bases[base_idx]; else g1a_neg(&b->p, &bases[base_idx]);
  } else {
    s->set[s->ptr].base_idx = base_idx;
    s->set[s->ptr].buck_idx = buck_idx;
    s->set[s->ptr].sign = sign;
    s->ptr++;
  }
  if (s->ptr == BATCH_SIZE) sched_execute(s, bases);
}

/* ------------------------------------------------------------------ msm_serial (msm.rs:338-423) */
static int window_size_for(size_t n) {
  if (n < 4) return 1;
  if (n < 32) return 3;
  /* ceil(ln n) computed exactly as f64::ln(n as u32).ceil() */
  return (int)ceil(log((double)(uint32_t)n));
}
int cpu_ref_window_size(size_t n) { return window_size_for(n); }

static void msm_serial(const uint8_t *reprs, const g1a_t *bases, size_t n, g1j_t *acc) {
  int c = window_size_for(n);
  uint8_t acc_or[32] = {0};
  for (size_t i = 0; i < n; i++)
    for (int b = 0; b < 32; b++) acc_or[b] |= reprs[32 * i + b];
  int max_byte = 32;
  while (max_byte > 0 && acc_or[max_byte - 1] == 0) max_byte--;
  if (max_byte == 0) return;
  int nw = max_byte * 8 / c + 1;
  size_t nb = (size_t)1 << (c - 1);
  /* bucket kinds: 0 None, 1 Affine, 2 Projective */
  int *kind = (int *)malloc(sizeof(int) * nb);
  g1a_t *ab = (g1a_t *)malloc(sizeof(g1a_t) * nb);
  g1j_t *pb = (g1j_t *)malloc(sizeof(g1j_t) * nb);
  for (int w = nw - 1; w >= 0; w--) {
    for (int k = 0; k < c; k++) g1j_dbl(acc, acc);
    memset(kind, 0, sizeof(int) * nb);
    for (size_t i = 0; i < n; i++) {
      int32_t d = get_booth_index(w, c, reprs + 32 * i);
      if (d == 0) continue;
      g1a_t q;
      if (d > 0) q = bases[i]; else g1a_neg(&q, &bases[i]);
      size_t bi = (size_t)(d > 0 ? d : -d) - 1;
      if (kind[bi] == 0) { kind[bi] = 1; ab[bi] = q; }
      else if (kind[bi] == 1) { g1j_t t; g1j_from_affine(&t, &ab[bi]); g1j_add_affine(&pb[bi], &t, &q); kind[bi] = 2; }
      else g1j_add_affine(&pb[bi], &pb[bi], &q);
    }
    g1j_t running;
    g1j_set_inf(&running);
    for (size_t bi = nb; bi-- > 0;) {
      if (kind[bi] == 1) g1j_add_affine(&running, &running, &ab[bi]);
      else if (kind[bi] == 2) g1j_add(&running, &pb[bi], &running);
      g1j_add(acc, acc, &running);
    }
  }
  free(kind); free(ab); free(pb);
}

/* ------------------------------------------------------------------ msm_parallel (msm.rs:430-455) */
typedef struct { const uint8_t *reprs; const g1a_t *bases; size_t n, chunk; g1j_t *results; } par_ctx_t;
static void par_chunk(void *v, int i) {
  par_ctx_t *c = (par_ctx_t *)v;
  size_t lo = (size_t)i * c->chunk;
  size_t hi = lo + c->chunk < c->n ? lo + c->chunk : c->n;
  g1j_set_inf(&c->results[i]);
  msm_serial(c->reprs + 32 * lo, c->bases + lo, hi - lo, &c->results[i]);
}
static void msm_parallel(const uint8_t *reprs, const g1a_t *bases, size_t n, g1j_t *out) {
  int nt = g_threads;
  g1j_set_inf(out);
  if (n > (size_t)nt) {
    size_t chunk = n / nt;
    size_t nchunks = (n + chunk - 1) / chunk;
    g1j_t *res = (g1j_t *)malloc(sizeof(g1j_t) * nchunks);
    par_ctx_t ctx = {reprs, bases, n, chunk, res};
    parallel_for((int)nchunks, nt, par_chunk, &ctx);
    for (size_t i = 0; i < nchunks; i++) g1j_add(out, out, &res[i]);
    free(res);
  } else {
    msm_serial(reprs, bases, n, out);
  }
}

/* ------------------------------------------------------------------ msm_best (msm.rs:460-532) */
typedef struct { const uint8_t *reprs; const g1a_t *bases; size_t n; int c; g1j_t *acc; } best_ctx_t;
static void best_window(void *v, int w) {
  best_ctx_t *b = (best_ctx_t *)v;
  int c = b->c;
  size_t nb = (size_t)1 << (c - 1);
  buck_jac_t *jb = (buck_jac_t *)calloc(nb, sizeof(buck_jac_t));
  schedule_t s;
  s.buckets = (buck_aff_t *)malloc(sizeof(buck_aff_t) * nb);
  for (size_t i = 0; i < nb; i++) s.buckets[i].inf = 1;
  s.ptr = 0;
  for (size_t i = 0; i < b->n; i++) {
    int32_t d = get_booth_index(w, c, b->reprs + 32 * i);
    if (d == 0) continue;
    int sign = d > 0;
    size_t bi = (size_t)(d > 0 ? d : -d) - 1;
    if (sched_contains(&s, bi)) {
      g1a_t q;
      if (sign) q = b->bases[i]; else g1a_neg(&q, &b->bases[i]);
      if (jb[bi].kind == 0) { g1j_from_affine(&jb[bi].p, &q); jb[bi].kind = 1; }
      else g1j_add_affine(&jb[bi].p, &jb[bi].p, &q);
    } else {
      sched_add(&s, b->bases, i, bi, sign);
    }
  }
  sched_execute(&s, b->bases);
  g1j_t running, acc;
  g1j_set_inf(&running);
  g1j_set_inf(&acc);
  for (size_t bi = nb; bi-- > 0;) {
    g1j_t t;
    if (jb[bi].kind) {
      if (!s.buckets[bi].inf) g1j_add_affine(&t, &jb[bi].p, &s.buckets[bi].p);
      else t = jb[bi].p;
    } else if (!s.buckets[bi].inf) {
      g1j_from_affine(&t, &s.buckets[bi].p);
    } else {
      g1j_set_inf(&t);
    }
    g1j_add(&running, &running, &t);
    g1j_add(&acc, &acc, &running);
  }
  for (int k = 0; k < c * w; k++) g1j_dbl(&acc, &acc);
  b->acc[w] = acc;
  free(jb);
  free(s.buckets);
}

static void msm_best_impl(const uint8_t *reprs, const g1a_t *bases, size_t n, g1j_t *out) {
  int c = window_size_for(n);
  g1j_set_inf(out);
  if (c < 10) { msm_parallel(reprs, bases, n, out); return; }
  /* drop identity bases (msm.rs:477-482) */
  uint8_t *r2 = (uint8_t *)malloc(32 * n);
  g1a_t *b2 = (g1a_t *)malloc(sizeof(g1a_t) * n);
  size_t m = 0;
  for (size_t i = 0; i < n; i++) {
    if (g1a_is_inf(&bases[i])) continue;
    memcpy(r2 + 32 * m, reprs + 32 * i, 32);
    b2[m++] = bases[i];
  }
  int nw = 255 / c + 1;
  g1j_t *acc = (g1j_t *)malloc(sizeof(g1j_t) * nw);
  best_ctx_t ctx = {r2, b2, m, c, acc};
  parallel_for(nw, g_threads, best_window, &ctx);
  for (int w = 0; w < nw; w++) g1j_add(out, out, &acc[w]);
  free(acc); free(r2); free(b2);
}

static uint8_t *scalars_to_reprs(const u64 *scalars_mont, size_t n) {
  uint8_t *r = (uint8_t *)malloc(32 * n + 1);
  for (size_t i = 0; i < n; i++) {
    fr_t a, c;
    memcpy(a.l, scalars_mont + 4 * i, 32);
    fr_from_mont(&c, &a);
    memcpy(r + 32 * i, c.l, 32);     /* little-endian host */
  }
  return r;
}

/* Exported MSM entry points: scalars n x 4 u64 Montgomery, bases n x 12 u64
 * (blst_p1_affine, Montgomery), result 18 u64 (blst_p1, Jacobian Montgomery). */
void cpu_ref_msm_best(const u64 *scalars, const u64 *bases, size_t n, u64 *out) {
  uint8_t *reprs = scalars_to_reprs(scalars, n);
  msm_best_impl(reprs, (const g1a_t *)bases, n, (g1j_t *)out);
  free(reprs);
}
void cpu_ref_msm_parallel(const u64 *scalars, const u64 *bases, size_t n, u64 *out) {
  uint8_t *reprs = scalars_to_reprs(scalars, n);
  msm_parallel(reprs, (const g1a_t *)bases, n, (g1j_t *)out);
  free(reprs);
}
/* msm_serial (msm.rs:338-423): *acc += the MSM, acc Jacobian in and out. */
void cpu_ref_msm_serial(const u64 *scalars, const u64 *bases, size_t n, u64 *acc) {
  uint8_t *reprs = scalars_to_reprs(scalars, n);
  msm_serial(reprs, (const g1a_t *)bases, n, (g1j_t *)acc);
  free(reprs);
}
/* msm_specific (kzg/msm.rs:174-201): filter zero scalars, then msm_best (the
 * blst branch below 2^19 non-zero terms is third-party code not present here;
 * it computes the same unique point). */
void cpu_ref_msm_specific(const u64 *scalars, const u64 *bases, size_t n, u64 *out) {
  u64 *s2 = (u64 *)malloc(32 * (n + 1));
  u64 *b2 = (u64 *)malloc(96 * (n + 1));
  size_t m = 0;
  for (size_t i = 0; i < n; i++) {
    const u64 *s = scalars + 4 * i;
    if (!(s[0] | s[1] | s[2] | s[3])) continue;
    memcpy(s2 + 4 * m, s, 32);
    memcpy(b2 + 12 * m, bases + 12 * i, 96);
    m++;
  }
  if (m == 0) g1j_set_inf((g1j_t *)out);
  else cpu_ref_msm_best(s2, b2, m, out);
  free(s2); free(b2);
}

void cpu_ref_g1_to_affine(const u64 *jac, u64 *aff) { g1j_to_affine((g1a_t *)aff, (const g1j_t *)jac); }

/* ------------------------------------------------------------------ FFT (fft.rs) */
static size_t bitreverse(size_t n, unsigned log_n) {
  size_t r = 0;
  for (unsigned i = 0; i < log_n; i++) { r = (r << 1) | (n & 1); n >>= 1; }
  return r;
}

void cpu_ref_compute_twiddles(const u64 *omega, unsigned log_n, u64 *out) {
  size_t half = (size_t)1 << (log_n - 1);
  fr_t w, om;
  memcpy(w.l, FR_ONE, 32);
  memcpy(om.l, omega, 32);
  for (size_t i = 0; i < half; i++) { memcpy(out + 4 * i, w.l, 32); fr_mul(&w, &w, &om); }
}

typedef struct { fr_t *a; size_t n, tc; const fr_t *tw; int depth; } rec_t;
static void recursive_butterfly(fr_t *a, size_t n, size_t tc, const fr_t *tw, int depth);
static void *rec_tramp(void *v) { rec_t *r = (rec_t *)v; recursive_butterfly(r->a, r->n, r->tc, r->tw, r->depth); return NULL; }
/* fft.rs:202-236 */
static void recursive_butterfly(fr_t *a, size_t n, size_t tc, const fr_t *tw, int depth) {
  if (n == 2) {
    fr_t t = a[1];
    a[1] = a[0];
    fr_add(&a[0], &a[0], &t);
    fr_sub(&a[1], &a[1], &t);
    return;
  }
  size_t h = n / 2;
  if (depth > 0) {
    pthread_t th;
    rec_t r = {a, h, tc * 2, tw, depth - 1};
    pthread_create(&th, NULL, rec_tramp, &r);
    recursive_butterfly(a + h, h, tc * 2, tw, depth - 1);
    pthread_join(th, NULL);
  } else {
    recursive_butterfly(a, h, tc * 2, tw, 0);
    recursive_butterfly(a + h, h, tc * 2, tw, 0);
  }
  fr_t t = a[h];
  a[h] = a[0];
  fr_add(&a[0], &a[0], &t);
  fr_sub(&a[h], &a[h], &t);
  for (size_t i = 1; i < h; i++) {
    fr_mul(&t, &a[h + i], &tw[i * tc]);
    a[h + i] = a[i];
    fr_add(&a[i], &a[i], &t);
    fr_sub(&a[h + i], &a[h + i], &t);
  }
}

static int log2_threads(void) { int l = 0; while ((1 << (l + 1)) <= g_threads) l++; return l; }

/* fft.rs:56-101: bit-reverse, then iterative DIT when log_n <= log2(threads),
 * else recursive DIT. */
void cpu_ref_best_fft_with_twiddles(u64 *a_raw, const u64 *tw_raw, unsigned log_n) {
  fr_t *a = (fr_t *)a_raw;
  const fr_t *tw = (const fr_t *)tw_raw;
  size_t n = (size_t)1 << log_n;
  for (size_t k = 0; k < n; k++) {
    size_t rk = bitreverse(k, log_n);
    if (k < rk) { fr_t t = a[k]; a[k] = a[rk]; a[rk] = t; }
  }
  if (log_n == 0) return;
  if ((int)log_n <= log2_threads()) {
    size_t chunk = 2, tchunk = n / 2;
    for (unsigned s = 0; s < log_n; s++) {
      for (size_t base = 0; base < n; base += chunk) {
        fr_t *l = a + base, *r = a + base + chunk / 2;
        fr_t t = r[0];
        r[0] = l[0];
        fr_add(&l[0], &l[0], &t);
        fr_sub(&r[0], &r[0], &t);
        for (size_t i = 1; i < chunk / 2; i++) {
          fr_mul(&t, &r[i], &tw[i * tchunk]);
          r[i] = l[i];
          fr_add(&l[i], &l[i], &t);
          fr_sub(&r[i], &r[i], &t);
        }
      }
      chunk *= 2;
      tchunk /= 2;
    }
  } else {
    recursive_butterfly(a, n, 1, tw, log2_threads());
  }
}

static inline void gs(fr_t *a, size_t i, size_t j, const fr_t *t) {   /* blst_fr_gs_bfly */
  fr_t s, d;
  fr_add(&s, &a[i], &a[j]);
  fr_sub(&d, &a[i], &a[j]);
  a[i] = s;
  fr_mul(&a[j], &d, t);
}
/* fft.rs:148-186 */
typedef struct { fr_t *a; size_t n, tc; const fr_t *tw; size_t nz; int depth; } dif_t;
static void recursive_dif_pruned(fr_t *a, size_t n, size_t tc, const fr_t *tw, size_t nz, int depth);
static void *dif_tramp(void *v) { dif_t *d = (dif_t *)v; recursive_dif_pruned(d->a, d->n, d->tc, d->tw, d->nz, d->depth); return NULL; }
static void recursive_dif_pruned(fr_t *a, size_t n, size_t tc, const fr_t *tw, size_t nz, int depth) {
  if (nz == 0) return;
  if (n == 2) { gs(a, 0, 1, &tw[0]); return; }
  size_t h = n / 2;
  if (nz <= h) {
    for (size_t i = 0; i < nz; i++) fr_mul(&a[i + h], &a[i], &tw[i * tc]);
  } else {
    for (size_t i = 0; i < h; i++) gs(a, i, i + h, &tw[i * tc]);
  }
  size_t cnz = nz < h ? nz : h;
  if (depth > 0) {
    pthread_t th;
    dif_t d = {a, h, tc * 2, tw, cnz, depth - 1};
    pthread_create(&th, NULL, dif_tramp, &d);
    recursive_dif_pruned(a + h, h, tc * 2, tw, cnz, depth - 1);
    pthread_join(th, NULL);
  } else {
    recursive_dif_pruned(a, h, tc * 2, tw, cnz, 0);
    recursive_dif_pruned(a + h, h, tc * 2, tw, cnz, 0);
  }
}
/* fft.rs:110-143 (the Fq branch) */
void cpu_ref_fft_coeff_to_extended(u64 *a_raw, const u64 *tw_raw, unsigned log_n, size_t n_real) {
  fr_t *a = (fr_t *)a_raw;
  size_t n = (size_t)1 << log_n;
  recursive_dif_pruned(a, n, 1, (const fr_t *)tw_raw, n_real, log2_threads());
  for (size_t k = 0; k < n; k++) {
    size_t rk = bitreverse(k, log_n);
    if (k < rk) { fr_t t = a[k]; a[k] = a[rk]; a[rk] = t; }
  }
}

/* ------------------------------------------------------------------ EvaluationDomain (domain.rs) */
void cpu_ref_root_of_unity(unsigned log_n, u64 *out) {       /* ROOT^(2^(S-log_n)) */
  fr_t w;
  memcpy(w.l, FR_ROOT, 32);
  for (unsigned i = log_n; i < 32; i++) fr_mul(&w, &w, &w);
  memcpy(out, w.l, 32);
}
void cpu_ref_fr_inv(const u64 *a, u64 *out) { fr_t r; fr_inv(&r, (const fr_t *)a); memcpy(out, r.l, 32); }
void cpu_ref_fr_mul(const u64 *a, const u64 *b, u64 *out) { fr_t r; fr_mul(&r, (const fr_t *)a, (const fr_t *)b); memcpy(out, r.l, 32); }

static void scale_all(fr_t *a, size_t n, const fr_t *s) { for (size_t i = 0; i < n; i++) fr_mul(&a[i], &a[i], s); }
static void distribute_powers_zeta(fr_t *a, size_t n, int into) {
  fr_t z, z2;
  memcpy(z.l, FR_ZETA, 32);
  fr_mul(&z2, &z, &z);
  const fr_t *p1 = into ? &z : &z2, *p2 = into ? &z2 : &z;
  for (size_t i = 0; i < n; i++) {
    int r = (int)(i % 3);
    if (r == 1) fr_mul(&a[i], &a[i], p1);
    else if (r == 2) fr_mul(&a[i], &a[i], p2);
  }
}

static fr_t *twiddles_for(unsigned log_n, int inverse) {
  fr_t w;
  cpu_ref_root_of_unity(log_n, w.l);
  if (inverse) fr_inv(&w, &w);
  fr_t *tw = (fr_t *)malloc(32 * (((size_t)1 << (log_n ? log_n - 1 : 0)) + 1));
  if (log_n) cpu_ref_compute_twiddles(w.l, log_n, tw->l);
  return tw;
}

/* forward DFT (coeff_to_lagrange, domain.rs:268-277) */
void cpu_ref_ntt(u64 *a, unsigned log_n) {
  fr_t *tw = twiddles_for(log_n, 0);
  cpu_ref_best_fft_with_twiddles(a, tw->l, log_n);
  free(tw);
}
/* lagrange_to_coeff (domain.rs:249-264 + ifft :389-396) */
void cpu_ref_intt(u64 *a, unsigned log_n) {
  fr_t *tw = twiddles_for(log_n, 1);
  cpu_ref_best_fft_with_twiddles(a, tw->l, log_n);
  free(tw);
  fr_t ninv;
  fr_from_u64(&ninv, (u64)1 << log_n);
  fr_inv(&ninv, &ninv);
  scale_all((fr_t *)a, (size_t)1 << log_n, &ninv);
}
/* coeff_to_extended (domain.rs:285-305): in = 2^k coefficients, out = 2^ext_k */
void cpu_ref_coeff_to_extended(const u64 *in, unsigned k, unsigned ext_k, u64 *out) {
  size_t n = (size_t)1 << k, N = (size_t)1 << ext_k;
  memcpy(out, in, 32 * n);
  memset(out + 4 * n, 0, 32 * (N - n));
  distribute_powers_zeta((fr_t *)out, n, 1);
  fr_t *tw = twiddles_for(ext_k, 0);
  cpu_ref_fft_coeff_to_extended(out, tw->l, ext_k, n);
  free(tw);
}
/* extended_to_coeff (domain.rs:312-328) */
void cpu_ref_extended_to_coeff(u64 *a, unsigned ext_k) {
  cpu_ref_intt(a, ext_k);
  distribute_powers_zeta((fr_t *)a, (size_t)1 << ext_k, 0);
}
/* divide_by_vanishing_poly (domain.rs:361-380) with t_evaluations (:104-140) */
void cpu_ref_divide_by_vanishing_poly(u64 *a, unsigned k, unsigned ext_k) {
  size_t m = (size_t)1 << (ext_k - k);
  fr_t *t = (fr_t *)malloc(32 * m);
  fr_t z, we, orig, step, cur, one;
  memcpy(z.l, FR_ZETA, 32);
  memcpy(one.l, FR_ONE, 32);
  cpu_ref_root_of_unity(ext_k, we.l);
  fr_pow_u64(&orig, &z, (u64)1 << k);
  fr_pow_u64(&step, &we, (u64)1 << k);
  cur = orig;
  for (size_t i = 0; i < m; i++) { t[i] = cur; fr_mul(&cur, &cur, &step); }
  for (size_t i = 0; i < m; i++) { fr_sub(&t[i], &t[i], &one); fr_inv(&t[i], &t[i]); }
  size_t N = (size_t)1 << ext_k;
  fr_t *v = (fr_t *)a;
  for (size_t i = 0; i < N; i++) fr_mul(&v[i], &v[i], &t[i & (m - 1)]);
  free(t);
}

/* ------------------------------------------------------------------ seeded streams */
static u64 splitmix64(u64 x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
/* Fq::from_uniform_bytes (fq.rs:712-736): a0 * R2 + a1 * R3 (Montgomery products) */
static void fr_from_uniform_words(fr_t *out, const u64 w[8]) {
  fr_t a0, a1, r2, r3, t0, t1;
  memcpy(a0.l, w, 32);
  memcpy(a1.l, w + 4, 32);
  memcpy(r2.l, FR_R2, 32);
  memcpy(r3.l, FR_R3, 32);
  fr_mul(&t0, &a0, &r2);
  fr_mul(&t1, &a1, &r3);
  fr_add(out, &t0, &t1);
}
typedef struct { u64 key, offset; size_t n; u64 *out; } seed_ctx_t;
static void seed_chunk(void *v, int c) {
  seed_ctx_t *s = (seed_ctx_t *)v;
  size_t lo = (size_t)c * 65536, hi = lo + 65536 < s->n ? lo + 65536 : s->n;
  for (size_t i = lo; i < hi; i++) {
    u64 w[8];
    u64 idx = s->offset + i;
    for (int j = 0; j < 8; j++) w[j] = splitmix64(s->key + idx * 8 + (u64)j);
    fr_from_uniform_words((fr_t *)(s->out + 4 * i), w);
  }
}
/* element i of stream (seed, stream) in Montgomery form */
void cpu_ref_fr_from_seed(u64 seed, u64 stream, u64 offset, size_t n, u64 *out) {
  u64 key = splitmix64(seed ^ (stream * 0xD1B54A32D192ED03ull));
  seed_ctx_t ctx = {key, offset, n, out};
  parallel_for((int)((n + 65535) / 65536), g_threads, seed_chunk, &ctx);
}

/* ------------------------------------------------------------------ SRS (params.rs:154-172) */
static const u64 G1X[6] = {0x5cb38790fd530c16ull, 0x7817fc679976fff5ull, 0x154f95c7143ba1c1ull,
                           0xf0ae6acdf3d0e747ull, 0xedce6ecc21dbf440ull, 0x120177419e0bfb75ull};
static const u64 G1Y[6] = {0xbaac93d50ce72271ull, 0x8c22631a7918fd8eull, 0xdd595f13570725ceull,
                           0x51ac582950405194ull, 0x0e1c8c3fad0059c0ull, 0x0bbc3efc5008a26aull};
void cpu_ref_g1_generator(u64 *out) { memcpy(out, G1X, 48); memcpy(out + 6, G1Y, 48); }

/* fixed-base comb: table[w][d] = d * 2^(8w) G, affine, w < 32, d < 256 */
static g1a_t *g_comb = NULL;
static pthread_mutex_t g_comb_mu = PTHREAD_MUTEX_INITIALIZER;
static void comb_build(void) {
  pthread_mutex_lock(&g_comb_mu);
  if (!g_comb) {
    g1a_t *tab = (g1a_t *)malloc(sizeof(g1a_t) * 32 * 256);
    g1j_t base;
    g1a_t g;
    memcpy(g.x.l, G1X, 48);
    memcpy(g.y.l, G1Y, 48);
    g1j_from_affine(&base, &g);
    for (int w = 0; w < 32; w++) {
      g1j_t acc;
      g1j_set_inf(&acc);
      memset(&tab[w * 256], 0, sizeof(g1a_t));
      for (int d = 1; d < 256; d++) {
        g1j_add(&acc, &acc, &base);
        g1j_to_affine(&tab[w * 256 + d], &acc);
      }
      for (int k = 0; k < 8; k++) g1j_dbl(&base, &base);
    }
    g_comb = tab;
  }
  pthread_mutex_unlock(&g_comb_mu);
}
static void fixed_base_mul(g1j_t *out, const fr_t *s_mont) {
  fr_t s;
  fr_from_mont(&s, s_mont);
  const uint8_t *b = (const uint8_t *)s.l;
  g1j_set_inf(out);
  for (int w = 0; w < 32; w++)
    if (b[w]) g1j_add_affine(out, out, &g_comb[w * 256 + b[w]]);
}
typedef struct { const fr_t *tau; size_t n; g1a_t *out; const fr_t *scalars; } srs_ctx_t;
static void srs_chunk(void *v, int c) {
  srs_ctx_t *s = (srs_ctx_t *)v;
  size_t lo = (size_t)c * 4096, hi = lo + 4096 < s->n ? lo + 4096 : s->n;
  size_t m = hi - lo;
  fr_t p;
  if (!s->scalars) fr_pow_u64(&p, s->tau, lo);
  g1j_t *j = (g1j_t *)malloc(sizeof(g1j_t) * m);
  for (size_t i = 0; i < m; i++) {
    if (s->scalars) { fixed_base_mul(&j[i], &s->scalars[lo + i]); continue; }
    fixed_base_mul(&j[i], &p);
    fr_mul(&p, &p, s->tau);
  }
  /* batch normalisation (Montgomery's trick) */
  fp_t *pre = (fp_t *)malloc(sizeof(fp_t) * m);
  fp_t acc;
  memcpy(acc.l, FP_ONE, 48);
  for (size_t i = 0; i < m; i++) { pre[i] = acc; if (!g1j_is_inf(&j[i])) fp_mul(&acc, &acc, &j[i].z); }
  fp_inv(&acc, &acc);
  for (size_t i = m; i-- > 0;) {
    if (g1j_is_inf(&j[i])) { memset(&s->out[lo + i], 0, sizeof(g1a_t)); continue; }
    fp_t zi, zi2, zi3;
    fp_mul(&zi, &acc, &pre[i]);
    fp_mul(&acc, &acc, &j[i].z);
    fp_sqr(&zi2, &zi);
    fp_mul(&zi3, &zi2, &zi);
    fp_mul(&s->out[lo + i].x, &j[i].x, &zi2);
    fp_mul(&s->out[lo + i].y, &j[i].y, &zi3);
  }
  free(pre);
  free(j);
}
/* g_i = [tau^i] G for i < n, affine Montgomery (blst_p1_affine layout) */
void cpu_ref_srs_bases(const u64 *tau_mont, size_t n, u64 *out) {
  comb_build();
  srs_ctx_t ctx = {(const fr_t *)tau_mont, n, (g1a_t *)out, NULL};
  parallel_for((int)((n + 4095) / 4096), g_threads, srs_chunk, &ctx);
}

/* sum_i s_i tau^i (Horner), Montgomery in/out */
void cpu_ref_poly_eval(const u64 *coeffs, size_t n, const u64 *x, u64 *out) {
  fr_t acc = {{0, 0, 0, 0}};
  for (size_t i = n; i-- > 0;) {
    fr_mul(&acc, &acc, (const fr_t *)x);
    fr_add(&acc, &acc, (const fr_t *)(coeffs + 4 * i));
  }
  memcpy(out, acc.l, 32);
}
/* [s]G for one scalar (Montgomery), Jacobian out */
void cpu_ref_g1_mul_generator(const u64 *s_mont, u64 *out_jac) {
  comb_build();
  fixed_base_mul((g1j_t *)out_jac, (const fr_t *)s_mont);
}

/* ------------------------------------------------------------------ scalar vectors
 * Restatements of the prover's element-wise Polynomial arithmetic (proofs/src/poly/mod.rs),
 * ff BatchInvert (used at permutation/prover.rs:148), the grand-product recurrence
 * (permutation/prover.rs:190-197), kate_division (utils/arithmetic.rs:106-125) and
 * powers (kzg/mod.rs:151-167). */
typedef struct { int op; const fr_t *a, *b; fr_t *out; size_t n; fr_t s; } vec_ctx_t;
#define VEC_CHUNK 16384
static void vec_binop_chunk(void *v, int c) {
  vec_ctx_t *x = (vec_ctx_t *)v;
  size_t lo = (size_t)c * VEC_CHUNK, hi = lo + VEC_CHUNK < x->n ? lo + VEC_CHUNK : x->n;
  for (size_t i = lo; i < hi; i++) {
    fr_t r;
    if (x->op == 0) fr_add(&r, &x->a[i], &x->b[i]);
    else if (x->op == 1) fr_sub(&r, &x->a[i], &x->b[i]);
    else fr_mul(&r, &x->a[i], &x->b[i]);
    x->out[i] = r;
  }
}
/* op 0: a + b, 1: a - b, 2: a * b */
void cpu_ref_fr_vec_op(int op, const u64 *a, const u64 *b, u64 *out, size_t n) {
  vec_ctx_t x = {op, (const fr_t *)a, (const fr_t *)b, (fr_t *)out, n, {{0, 0, 0, 0}}};
  parallel_for((int)((n + VEC_CHUNK - 1) / VEC_CHUNK), g_threads, vec_binop_chunk, &x);
}
static void vec_scalar_chunk(void *v, int c) {
  vec_ctx_t *x = (vec_ctx_t *)v;
  size_t lo = (size_t)c * VEC_CHUNK, hi = lo + VEC_CHUNK < x->n ? lo + VEC_CHUNK : x->n;
  for (size_t i = lo; i < hi; i++) {
    fr_t r;
    if (x->op == 0) fr_mul(&r, &x->a[i], &x->s);
    else if (x->op == 1) fr_add(&r, &x->a[i], &x->s);
    else { fr_mul(&r, &x->a[i], &x->s); fr_add(&r, &x->out[i], &r); }
    x->out[i] = r;
  }
}
/* op 0: out = a * s, 1: out = a + s, 2: out = out + a * s */
void cpu_ref_fr_vec_scalar(int op, const u64 *a, const u64 *s, u64 *out, size_t n) {
  vec_ctx_t x = {op, (const fr_t *)a, NULL, (fr_t *)out, n, *(const fr_t *)s};
  parallel_for((int)((n + VEC_CHUNK - 1) / VEC_CHUNK), g_threads, vec_scalar_chunk, &x);
}
/* Montgomery's trick over the whole vector, zeros skipped (ff::BatchInvert) */
void cpu_ref_fr_batch_invert(u64 *a_raw, size_t n) {
  fr_t *a = (fr_t *)a_raw;
  fr_t *pre = (fr_t *)malloc(sizeof(fr_t) * (n ? n : 1));
  fr_t acc;
  memcpy(acc.l, FR_ONE, 32);
  for (size_t i = 0; i < n; i++) { pre[i] = acc; if (!fr_is_zero(&a[i])) fr_mul(&acc, &acc, &a[i]); }
  fr_inv(&acc, &acc);
  for (size_t i = n; i-- > 0;) {
    if (fr_is_zero(&a[i])) continue;
    fr_t t;
    fr_mul(&t, &acc, &pre[i]);
    fr_mul(&acc, &acc, &a[i]);
    a[i] = t;
  }
  free(pre);
}
/* out_i = a_0 * ... * a_i */
void cpu_ref_fr_prefix_product(const u64 *a_raw, u64 *out_raw, size_t n) {
  const fr_t *a = (const fr_t *)a_raw;
  fr_t *out = (fr_t *)out_raw;
  fr_t acc;
  memcpy(acc.l, FR_ONE, 32);
  for (size_t i = 0; i < n; i++) { fr_mul(&acc, &acc, &a[i]); out[i] = acc; }
}
/* kate_division: b <- -b; walking down from the top, q = r - tmp, tmp = q * b */
void cpu_ref_fr_kate_div(const u64 *a_raw, size_t n, const u64 *b_raw, u64 *q_raw) {
  const fr_t *a = (const fr_t *)a_raw;
  fr_t *q = (fr_t *)q_raw;
  fr_t zero = {{0, 0, 0, 0}}, b, tmp = zero;
  fr_sub(&b, &zero, (const fr_t *)b_raw);
  for (size_t i = n - 1; i-- > 0;) {
    fr_t lead;
    fr_sub(&lead, &a[i + 1], &tmp);
    q[i] = lead;
    fr_mul(&tmp, &lead, &b);
  }
}
/* out_i = x^i */
void cpu_ref_fr_powers(const u64 *x, size_t n, u64 *out_raw) {
  fr_t *out = (fr_t *)out_raw;
  fr_t acc;
  memcpy(acc.l, FR_ONE, 32);
  for (size_t i = 0; i < n; i++) { out[i] = acc; fr_mul(&acc, &acc, (const fr_t *)x); }
}
/* [s_i]G for n scalars (the Lagrange bases of unsafe_setup, params.rs:174-186), affine out */
/* suffix_sum (params.rs:47-57): out_i = sum_(j >= i) in_j, affine; one running Jacobian
 * sum from the end, then all points normalised with one inversion (Montgomery's trick). */
void cpu_ref_g1_suffix_sum(const u64 *in, size_t n, u64 *out) {
  g1j_t *acc = (g1j_t *)malloc(sizeof(g1j_t) * (n + 1));
  fp_t *pre = (fp_t *)malloc(sizeof(fp_t) * (n + 1));
  g1j_t run;
  g1j_set_inf(&run);
  for (size_t i = n; i-- > 0;) {
    g1j_add_affine(&run, &run, (const g1a_t *)(in + 12 * i));
    acc[i] = run;
  }
  fp_t p;
  memcpy(p.l, FP_ONE, 48);
  for (size_t i = 0; i < n; i++) {          /* prefix products of the non-zero z */
    pre[i] = p;
    if (!g1j_is_inf(&acc[i])) fp_mul(&p, &p, &acc[i].z);
  }
  fp_t inv;
  fp_inv(&inv, &p);
  for (size_t i = n; i-- > 0;) {
    g1a_t *r = (g1a_t *)(out + 12 * i);
    if (g1j_is_inf(&acc[i])) { memset(r, 0, sizeof *r); continue; }
    fp_t zi, zi2, zi3;
    fp_mul(&zi, &inv, &pre[i]);               /* 1 / z_i */
    fp_mul(&inv, &inv, &acc[i].z);
    fp_sqr(&zi2, &zi);
    fp_mul(&zi3, &zi2, &zi);
    fp_mul(&r->x, &acc[i].x, &zi2);
    fp_mul(&r->y, &acc[i].y, &zi3);
  }
  free(acc);
  free(pre);
}

/* ------------------------------------------------------------------ LogUp (logup/prover.rs) */
typedef struct { fr_t v; size_t idx; } keyed_t;
static int fr_cmp(const fr_t *a, const fr_t *b) {
  for (int i = 3; i >= 0; i--)
    if (a->l[i] != b->l[i]) return a->l[i] < b->l[i] ? -1 : 1;
  return 0;
}
static int keyed_cmp(const void *x, const void *y) {
  const keyed_t *a = (const keyed_t *)x, *b = (const keyed_t *)y;
  int c = fr_cmp(&a->v, &b->v);
  return c ? c : (a->idx < b->idx ? -1 : a->idx > b->idx);
}
/* compute_multiplicities (:411-477): counts of the usable table rows' values among the
 * selected usable input rows, over the table counts; blinding values on the other rows.
 * Returns 0, or 1 + (c n + i) for the first selected input value (column c, row i; columns
 * outer, as the reference iterates) missing from the table, where the reference panics. */
long long cpu_ref_logup_multiplicities(const u64 *sel, const u64 *const *inputs, size_t ninputs, const u64 *table,
                                       size_t n, size_t usable, const u64 *blinding, u64 *m_out) {
  keyed_t *k = (keyed_t *)malloc(sizeof(keyed_t) * (usable + 1));
  size_t *run_of = (size_t *)malloc(sizeof(size_t) * (usable + 1));   /* sorted position -> run */
  size_t *pos_of = (size_t *)malloc(sizeof(size_t) * (usable + 1));   /* table row -> sorted position */
  u64 *tcnt = (u64 *)calloc(usable + 1, 8), *icnt = (u64 *)calloc(usable + 1, 8);
  for (size_t i = 0; i < usable; i++) { memcpy(k[i].v.l, table + 4 * i, 32); k[i].idx = i; }
  qsort(k, usable, sizeof(keyed_t), keyed_cmp);
  size_t runs = 0;
  for (size_t j = 0; j < usable; j++) {
    if (j > 0 && fr_cmp(&k[j].v, &k[j - 1].v) != 0) runs++;
    run_of[j] = runs;
    pos_of[k[j].idx] = j;
    tcnt[runs]++;
  }
  long long bad = 0;
  for (size_t c = 0; c < ninputs && !bad; c++)
    for (size_t i = 0; i < usable; i++) {
      const u64 *s = sel + 4 * i;
      if (!(s[0] | s[1] | s[2] | s[3])) continue;
      fr_t v;
      memcpy(v.l, inputs[c] + 4 * i, 32);
      size_t lo = 0, hi = usable;   /* first position with value >= v */
      while (lo < hi) {
        size_t mid = (lo + hi) / 2;
        if (fr_cmp(&k[mid].v, &v) < 0) lo = mid + 1; else hi = mid;
      }
      if (lo == usable || fr_cmp(&k[lo].v, &v) != 0) { bad = 1 + (long long)(c * n + i); break; }
      icnt[run_of[lo]]++;
    }
  if (!bad) {
    u64 *den = (u64 *)malloc(32 * (n + 1));
    for (size_t i = 0; i < n; i++) {
      if (i < usable) {
        size_t r = run_of[pos_of[i]];
        fr_from_u64((fr_t *)(m_out + 4 * i), icnt[r]);
        fr_from_u64((fr_t *)(den + 4 * i), tcnt[r]);
      } else {
        memcpy(m_out + 4 * i, blinding + 4 * (i - usable), 32);
        memcpy(den + 4 * i, FR_ONE, 32);
      }
    }
    cpu_ref_fr_batch_invert(den, n);
    for (size_t i = 0; i < usable; i++) fr_mul((fr_t *)(m_out + 4 * i), (fr_t *)(m_out + 4 * i), (fr_t *)(den + 4 * i));
    free(den);
  }
  free(k); free(run_of); free(pos_of); free(tcnt); free(icnt);
  return bad;
}

/* compute_logderivative (:189-310): helpers per chunk (sums of batch-inverted input
 * denominators), the row log-derivatives and the aggregator with its blinding rows. */
void cpu_ref_logup_logderivative(size_t n, size_t bf, const u64 *sel, const u64 *m, const u64 *table,
                                 const u64 *const *inputs, size_t ninputs, const uint32_t *chunk_sizes, size_t nchunks,
                                 const u64 *beta, const u64 *blinding, u64 *const *helpers, u64 *z) {
  (void)ninputs;
  u64 *tden = (u64 *)malloc(32 * n), *den = (u64 *)malloc(32 * n), *logd = (u64 *)malloc(32 * n);
  for (size_t i = 0; i < n; i++) fr_add((fr_t *)(tden + 4 * i), (const fr_t *)beta, (const fr_t *)(table + 4 * i));
  cpu_ref_fr_batch_invert(tden, n);
  size_t col = 0;
  for (size_t k = 0; k < nchunks; k++) {
    fr_t *h = (fr_t *)helpers[k];
    memset(h, 0, 32 * n);
    for (uint32_t c = 0; c < chunk_sizes[k]; c++, col++) {
      for (size_t i = 0; i < n; i++) fr_add((fr_t *)(den + 4 * i), (const fr_t *)beta, (const fr_t *)(inputs[col] + 4 * i));
      cpu_ref_fr_batch_invert(den, n);
      for (size_t i = 0; i < n; i++) fr_add(&h[i], &h[i], (const fr_t *)(den + 4 * i));
    }
  }
  for (size_t i = 0; i < n; i++) {
    fr_t sum, a, b;
    memset(&sum, 0, 32);
    for (size_t k = 0; k < nchunks; k++) fr_add(&sum, &sum, (const fr_t *)(helpers[k] + 4 * i));
    fr_mul(&a, (const fr_t *)(sel + 4 * i), &sum);
    fr_mul(&b, (const fr_t *)(m + 4 * i), (const fr_t *)(tden + 4 * i));
    fr_sub((fr_t *)(logd + 4 * i), &a, &b);
  }
  fr_t state;
  memset(&state, 0, 32);
  memset(z, 0, 32);
  for (size_t j = 0; j + 1 < n - bf; j++) {
    fr_add(&state, &state, (const fr_t *)(logd + 4 * j));
    memcpy(z + 4 * (j + 1), state.l, 32);
  }
  for (size_t t = 0; t < bf; t++) memcpy(z + 4 * (n - bf + t), blinding + 4 * t, 32);
  free(tden); free(den); free(logd);
}

/* Polynomial::to_delta / to_double_delta (poly/mod.rs:290-335). */
void cpu_ref_fr_to_delta(const u64 *a_raw, size_t n, int order, u64 *out_raw) {
  const fr_t *a = (const fr_t *)a_raw;
  fr_t *o = (fr_t *)out_raw;
  for (size_t i = 0; i < n; i++) {
    fr_t r = a[i];
    if (i >= 1) {
      fr_sub(&r, &r, &a[i - 1]);
      if (order == 2) {
        fr_sub(&r, &r, &a[i - 1]);
        if (i >= 2) fr_add(&r, &r, &a[i - 2]);
      }
    }
    o[i] = r;
  }
}

void cpu_ref_g1_mul_generator_many(const u64 *s_mont, size_t n, u64 *out_aff) {
  comb_build();
  srs_ctx_t ctx = {NULL, n, (g1a_t *)out_aff, (const fr_t *)s_mont};
  parallel_for((int)((n + 4095) / 4096), g_threads, srs_chunk, &ctx);
}
