// Reads commands "op n" and n records of 8-word canonical blst fr values (LE u32), writes
// the results as 8-word canonical values: the host build of the radix-2^29 field of the
// NTT passes, for comparison with Python integers.
//   op 0: to_std(from_std(a))                      (round trip)
//   op 1: to_std(mul(from_std(a), from_std(b)))    (Montgomery product)
//   op 2: to_std(reduce_small(from_std(a)))
//   op 3: to_std(add(ra, mul(rb, ra))), ra = reduce_small(from_std(a)), likewise rb
//   op 4: to_std(sub_kr<2>(ra, mul(rb, ra)))   (the product is below 2 r: b < 2 r holds)
//   op 5: 11 butterfly stages x <- x + w y, y <- x - w y + 2 r from reduced inputs,
//         twiddle w = reduce_small(from_std(b)) fully reduced below r; outputs x, y
//   op 6: one register pair of stages as the NTT runs it, the first stage's sums left
//         lazy (add_lazy, sub3r_lazy), from inputs raised by 22 r (the largest values a
//         stage pair sees): x = (a, b, b, a) + 22 r, w1 = b, w2 = a below r; outputs the
//         four results (32 words)
#include <stdio.h>
#include <stdint.h>
#include <vector>

#include "../../midnight-zk_amd/csrc/fr29.hpp"

using namespace mzk;

static fr29 below_r(const uint32_t* s) {   // a representative below r
  fr29 v = fr29_reduce_small(fr29_from_std(s));
  uint32_t c[8];
  fr29_to_std(v, c);   // canonical x 2^256
  fr29 o = fr29_from_std(c);
  // from_std of a canonical value is 32 x, reduce it fully: compare-free, reduce twice
  o = fr29_reduce_small(o);
  return o;
}

int main() {
  int op;
  unsigned n;
  while (scanf("%d %u", &op, &n) == 2) {
    getchar();
    std::vector<uint32_t> a(8 * n), b(8 * n);
    if (fread(a.data(), 4, 8 * n, stdin) != 8 * n) return 1;
    if (fread(b.data(), 4, 8 * n, stdin) != 8 * n) return 1;
    for (unsigned i = 0; i < n; i++) {
      const uint32_t* x = &a[8 * i];
      const uint32_t* y = &b[8 * i];
      uint32_t out[32] = {0};
      int words = 8;
      if (op == 0) {
        fr29_to_std(fr29_from_std(x), out);
      } else if (op == 1) {
        fr29_to_std(fr29_mul(fr29_from_std(x), fr29_from_std(y)), out);
      } else if (op == 2) {
        fr29_to_std(fr29_reduce_small(fr29_from_std(x)), out);
      } else if (op == 3 || op == 4) {
        const fr29 ra = fr29_reduce_small(fr29_from_std(x)), rb = fr29_reduce_small(fr29_from_std(y));
        const fr29 t = fr29_mul(rb, ra);
        fr29_to_std(op == 3 ? fr29_add(ra, t) : fr29_sub_kr<2>(ra, t), out);
      } else if (op == 6) {
        fr29 r22 = {{0}};   // 22 r, normalised
        const fr29 rr = {{r29c::R(0), r29c::R(1), r29c::R(2), r29c::R(3), r29c::R(4), r29c::R(5), r29c::R(6),
                          r29c::R(7), r29c::R(8)}};
        for (int k = 0; k < 22; k++) r22 = fr29_add(r22, rr);
        const fr29 A = fr29_add(fr29_reduce_small(fr29_from_std(x)), r22);
        const fr29 B = fr29_add(fr29_reduce_small(fr29_from_std(y)), r22);
        const fr29 w1 = below_r(y), w2 = below_r(x);
        const fr29 t1 = fr29_mul(B, w1), t3 = fr29_mul(A, w1);
        const fr29 y0 = fr29_add_lazy(A, t1), y1 = fr29_sub_kr<2>(A, t1);
        const fr29 y2 = fr29_add_lazy(B, t3), y3 = fr29_sub3r_lazy(B, t3);
        const fr29 t2 = fr29_mul(y2, w2), t4 = fr29_mul(y3, w2);
        fr29_to_std(fr29_add(y0, t2), out);
        fr29_to_std(fr29_sub_kr<2>(y0, t2), out + 8);
        fr29_to_std(fr29_add(y1, t4), out + 16);
        fr29_to_std(fr29_sub_kr<2>(y1, t4), out + 24);
        words = 32;
      } else {
        fr29 X = fr29_reduce_small(fr29_from_std(x)), Y = fr29_reduce_small(fr29_from_std(y));
        const fr29 w = below_r(y);
        for (int s = 0; s < 11; s++) {
          const fr29 t = fr29_mul(w, Y);
          const fr29 nx = fr29_add(X, t);
          Y = fr29_sub_kr<2>(X, t);
          X = nx;
        }
        fr29_to_std(X, out);
        fr29_to_std(Y, out + 8);
        words = 16;
      }
      fwrite(out, 4, words, stdout);
    }
  }
  return 0;
}
