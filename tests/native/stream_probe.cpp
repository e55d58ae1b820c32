// Test helper: evaluates the product's stream header (pyhea_amd/csrc/
// hea_stream.h) on the host so its Philox and permutation can be compared
// with the oracle's independent implementation.
//   stream_probe philox c0 c1 c2 c3 k0 k1      -> 4 words
//   stream_probe pairs seed iter phase slot n count -> count lines "s1 s2"
//   stream_probe perm seed iter phase slot n c0 c1 ... -> n species
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "hea_stream.h"

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  if (!strcmp(argv[1], "philox") && argc == 8) {
    hea_u32x4 c = {(uint32_t)strtoul(argv[2], 0, 0), (uint32_t)strtoul(argv[3], 0, 0),
                   (uint32_t)strtoul(argv[4], 0, 0), (uint32_t)strtoul(argv[5], 0, 0)};
    hea_u32x4 o = hea_philox4x32_10(c, (uint32_t)strtoul(argv[6], 0, 0),
                                    (uint32_t)strtoul(argv[7], 0, 0));
    printf("%u %u %u %u\n", o.x, o.y, o.z, o.w);
    return 0;
  }
  if (!strcmp(argv[1], "pairs") && argc == 8) {
    uint64_t seed = strtoull(argv[2], 0, 0);
    uint32_t it = strtoul(argv[3], 0, 0), ph = strtoul(argv[4], 0, 0), sl = strtoul(argv[5], 0, 0);
    uint32_t n = strtoul(argv[6], 0, 0), cnt = strtoul(argv[7], 0, 0);
    for (uint32_t p = 0; p < cnt; ++p) {
      uint32_t a, b;
      hea_stream_pair(seed, it, ph, sl, p, n, &a, &b);
      printf("%u %u\n", a, b);
    }
    return 0;
  }
  if (!strcmp(argv[1], "perm") && argc >= 8) {
    uint64_t seed = strtoull(argv[2], 0, 0);
    uint32_t it = strtoul(argv[3], 0, 0), ph = strtoul(argv[4], 0, 0), sl = strtoul(argv[5], 0, 0);
    uint32_t n = strtoul(argv[6], 0, 0);
    int k = argc - 7;
    int32_t prefix[17] = {0};
    for (int a = 0; a < k; ++a) prefix[a + 1] = prefix[a] + atoi(argv[7 + a]);
    hea_perm_key key = hea_perm_make_key(seed, it, ph, sl, n);
    for (uint32_t i = 0; i < n; ++i)
      printf("%u\n", hea_species_of_rank(prefix, k, hea_perm_apply(&key, i)));
    return 0;
  }
  return 2;
}
