// TEST INFRASTRUCTURE ONLY: keyed random streams of the CPU oracle, shared
// by the oracle (hcs_oracle.cpp) and the reference harness (ref_harness.cpp).
//
// Written from the published Philox4x32-10 definition (Salmon, Moraes, Dror,
// Shaw, SC'11), independently of the product's pyhea_amd/csrc/hea_stream.h,
// and pinned by the Random123 known-answer vectors in tests/.
//   word(p) = Philox4x32-10(key=seed, ctr=(p>>2, slot, iter, phase))[p&3] & 0x7fffffff
//   draw p  = (word(2p) % N, word(2p+1) % N)
// The permutation is a 6-round balanced Feistel network on 2h bits whose
// round keys are stream words 2^33 .. 2^33+5, cycle-walked into [0, N).
#pragma once
#include <stdint.h>
#include <vector>

namespace hea_oracle {

struct U4 {
  uint32_t v[4];
};

U4 philox(U4 c, uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    uint64_t a = (uint64_t)0xD2511F53u * c.v[0];
    uint64_t b = (uint64_t)0xCD9E8D57u * c.v[2];
    U4 n;
    n.v[0] = (uint32_t)(b >> 32) ^ c.v[1] ^ k0;
    n.v[1] = (uint32_t)b;
    n.v[2] = (uint32_t)(a >> 32) ^ c.v[3] ^ k1;
    n.v[3] = (uint32_t)a;
    c = n;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

struct Stream {
  uint64_t seed;
  uint32_t iter, phase, slot;
  uint64_t pos;  // next 31-bit word
  uint32_t next() {
    U4 c = {{(uint32_t)(pos >> 2), slot, iter, phase}};
    U4 o = philox(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    uint32_t w = o.v[pos & 3] & 0x7fffffffu;
    ++pos;
    return w;
  }
};

uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

// Keyed permutation of the sorted species multiset into a lattice.
void permute_lattice(uint64_t seed, uint32_t iter, uint32_t phase,
                     uint32_t slot, const int32_t* counts, int K, int N,
                     uint8_t* out) {
  uint32_t key[6];
  {
    Stream s{seed, iter, phase, slot, (uint64_t)0x80000000u << 2};
    for (int i = 0; i < 6; ++i) key[i] = s.next();
  }
  uint32_t bits = 2;
  while ((1u << bits) < (uint32_t)N) ++bits;
  const uint32_t h = (bits + 1) / 2, mask = (1u << h) - 1u;
  std::vector<int32_t> prefix(K + 1, 0);
  for (int a = 0; a < K; ++a) prefix[a + 1] = prefix[a] + counts[a];
  for (int i = 0; i < N; ++i) {
    uint32_t x = (uint32_t)i;
    do {
      uint32_t l = x >> h, r = x & mask;
      for (int q = 0; q < 6; ++q) {
        uint32_t f = mix32(r ^ key[q]) & mask;
        uint32_t nl = r;
        r = l ^ f;
        l = nl;
      }
      x = (l << h) | r;
    } while (x >= (uint32_t)N);
    int sp = 0;
    while (sp + 1 < K && (int32_t)x >= prefix[sp + 1]) ++sp;
    out[i] = (uint8_t)sp;
  }
}

}  // namespace hea_oracle
