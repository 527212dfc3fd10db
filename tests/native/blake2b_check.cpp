// Reads records "keylen perslen msglen" + key + personal + message from stdin and
// writes each 64-byte digest of the host BLAKE2b (csrc/blake2b.hpp); the message is fed
// in pieces of varying length, and a digest is also taken half-way to check that
// digest() leaves the state usable.
#include <stdio.h>
#include <stdint.h>
#include <vector>

#include "../../midnight-zk_amd/csrc/blake2b.hpp"

int main() {
  unsigned kl, pl, ml;
  while (scanf("%u %u %u", &kl, &pl, &ml) == 3) {
    getchar();   // the newline after the header
    std::vector<uint8_t> key(kl), pers(pl), msg(ml);
    if (kl && fread(key.data(), 1, kl, stdin) != kl) return 1;
    if (pl && fread(pers.data(), 1, pl, stdin) != pl) return 1;
    if (ml && fread(msg.data(), 1, ml, stdin) != ml) return 1;
    mzk::host::Blake2b h(key.data(), kl, pl ? pers.data() : nullptr);
    uint8_t mid[64], out[64];
    size_t off = 0, step = 1;
    while (off < ml) {
      const size_t take = step < ml - off ? step : ml - off;
      h.update(msg.data() + off, take);
      off += take;
      step = step * 3 + 1;
      if (off >= ml / 2 && off < ml) h.digest(mid);
    }
    h.digest(out);
    fwrite(out, 1, 64, stdout);
  }
  return 0;
}
