// Host harness for the radix-2^28 field and XYZZ code of the MSM kernels
// (midnight-zk_amd/csrc/fp28.hpp, g1_28.hpp), driven by tests/test_fp28_host.py.
//
// stdin: a command word, then binary little-endian u64 data; stdout: binary results.
//   mul  n | n pairs of blst fp (6 u64 each)      -> n products (6 u64 each)
//   sqr  n | n fp                                 -> n squares
//   madd n | n affine points (12 u64) + n signs (u64) -> 12 u64 Jacobian-free XYZZ
//                                                    (X, Y, ZZ, ZZZ canonical, 24 u64)
//   add  n | n affine points, two chains summed by madd then added with g1x28_add,
//            and the sum doubled with g1x28_dbl     -> 2 XYZZ (48 u64)
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../midnight-zk_amd/csrc/g1_28.hpp"

using namespace mzk;

static std::vector<uint64_t> read_u64(size_t n) {
  std::vector<uint64_t> v(n);
  if (n && fread(v.data(), 8, n, stdin) != n) {
    fprintf(stderr, "short input\n");
    std::exit(2);
  }
  return v;
}
static void write_u32(const uint32_t* p, size_t n) { fwrite(p, 4, n, stdout); }

int main() {
  char cmd[16] = {0};
  uint64_t n = 0;
  if (scanf("%15s %lu", cmd, &n) != 2) return 2;
  getchar();   // the newline before the binary payload
  const std::string c(cmd);
  if (c == "mul" || c == "sqr") {
    const int per = c == "mul" ? 12 : 6;
    auto in = read_u64(n * per);
    for (uint64_t i = 0; i < n; i++) {
      const uint32_t* a = (const uint32_t*)&in[i * per];
      const uint32_t* b = c == "mul" ? a + 12 : a;
      f28 r = c == "mul" ? f28_mul(f28_from_std(a), f28_from_std(b)) : f28_sqr(f28_from_std(a));
      uint32_t out[12];
      f28_to_std(r, out);
      write_u32(out, 12);
    }
    return 0;
  }
  if (c == "madd" || c == "add") {
    auto pts = read_u64(n * 12);
    auto sg = read_u64(n);
    g1x28 acc[2] = {g1x28_inf(), g1x28_inf()};
    for (uint64_t i = 0; i < n; i++) {
      const uint32_t* q = (const uint32_t*)&pts[i * 12];
      uint32_t any = 0;
      for (int j = 0; j < 24; j++) any |= q[j];
      if (!any) continue;
      g1x28& a = acc[c == "add" ? (i & 1) : 0];
      a = g1x28_madd(a, f28_from_std(q), f28_from_std(q + 12), sg[i] != 0);
    }
    uint32_t out[48];
    if (c == "madd") {
      g1x28_to_std(acc[0], out);
      write_u32(out, 48);
    } else {
      const g1x28 s = g1x28_add(acc[0], acc[1]);
      g1x28_to_std(s, out);
      write_u32(out, 48);
      g1x28_to_std(g1x28_dbl(s), out);
      write_u32(out, 48);
    }
    return 0;
  }
  return 2;
}
