// Keyed random streams shared by every stage of the SRO lattice search.
//
// The reference engine draws proposals from an unseeded global std::rand()
// (pyhea/cpp/accelerate.cpp:649-654) and permutations from
// std::mt19937(std::random_device()) (accelerate.cpp:241-246), so no two runs
// agree.  Here every draw is a pure function of (seed, iter, phase, slot,
// position), which lets thousands of GPU lanes generate the stream in
// parallel and lets a CPU run replay it exactly.
//
//   word(p)  = Philox4x32-10(key = seed, ctr = (p >> 2, slot, iter, phase))[p & 3]
//              & 0x7fffffff                      (31 bits, like RAND_MAX)
//   pair(p)  = (word(2p) % N, word(2p+1) % N)    (atom1, atom2 of draw p)
//
// Philox4x32-10 is the counter-based generator of Salmon, Moraes, Dror and
// Shaw, "Parallel random numbers: as easy as 1, 2, 3" (SC'11).
//
// The permutation for (seed, iter, phase, slot) is a balanced Feistel
// network on 2h bits (2^(2h) >= N) with cycle walking, so site i receives
// sorted_multiset[perm(i)] independently of every other site.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define HEA_HD __host__ __device__ __forceinline__
#else
#define HEA_HD static inline
#endif

// Stream phases (same values as the public header pyhea_hcs.h).
#ifndef HEA_PHASE_INIT
#define HEA_PHASE_INIT 0  // initial population        (accelerate.cpp:737)
#define HEA_PHASE_NEW 1   // per-iteration new lattices (accelerate.cpp:743)
#define HEA_PHASE_LS1 2   // local search I proposals   (accelerate.cpp:745)
#define HEA_PHASE_LS2 3   // local search II proposals  (accelerate.cpp:749)
#endif

struct hea_u32x4 {
  uint32_t x, y, z, w;
};

HEA_HD hea_u32x4 hea_philox4x32_10(hea_u32x4 c, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)M0 * c.x;
    const uint64_t p1 = (uint64_t)M1 * c.z;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    hea_u32x4 n;
    n.x = hi1 ^ c.y ^ k0;
    n.y = lo1;
    n.z = hi0 ^ c.w ^ k1;
    n.w = lo0;
    c = n;
    k0 += W0;
    k1 += W1;
  }
  return c;
}

// The four 31-bit words of block `blk` of stream (seed, iter, phase, slot).
HEA_HD hea_u32x4 hea_stream_block(uint64_t seed, uint32_t iter, uint32_t phase,
                                  uint32_t slot, uint32_t blk) {
  hea_u32x4 c = {blk, slot, iter, phase};
  hea_u32x4 o = hea_philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  o.x &= 0x7fffffffu;
  o.y &= 0x7fffffffu;
  o.z &= 0x7fffffffu;
  o.w &= 0x7fffffffu;
  return o;
}

// Proposal draw p of a stream: both site indices in [0, n).
HEA_HD void hea_stream_pair(uint64_t seed, uint32_t iter, uint32_t phase,
                            uint32_t slot, uint32_t p, uint32_t n,
                            uint32_t* s1, uint32_t* s2) {
  // words 2p and 2p+1 always lie in the same Philox block
  hea_u32x4 o = hea_stream_block(seed, iter, phase, slot, p >> 1);
  const uint32_t w1 = (p & 1u) ? o.z : o.x;
  const uint32_t w2 = (p & 1u) ? o.w : o.y;
  *s1 = w1 % n;
  *s2 = w2 % n;
}

// 32-bit avalanche mixer ("lowbias32", C. Wellons, hash-prospector).
HEA_HD uint32_t hea_mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

#define HEA_FEISTEL_ROUNDS 6

struct hea_perm_key {
  uint32_t k[HEA_FEISTEL_ROUNDS];
  uint32_t half_bits;  // h: each Feistel half has h bits, domain 2^(2h) >= n
  uint32_t n;
};

HEA_HD hea_perm_key hea_perm_make_key(uint64_t seed, uint32_t iter,
                                      uint32_t phase, uint32_t slot,
                                      uint32_t n) {
  hea_perm_key pk;
  // Round keys use stream blocks 0x80000000.. so they never overlap the
  // proposal words of a stream with the same (iter, phase, slot).
  hea_u32x4 a = hea_stream_block(seed, iter, phase, slot, 0x80000000u);
  hea_u32x4 b = hea_stream_block(seed, iter, phase, slot, 0x80000001u);
  pk.k[0] = a.x;
  pk.k[1] = a.y;
  pk.k[2] = a.z;
  pk.k[3] = a.w;
  pk.k[4] = b.x;
  pk.k[5] = b.y;
  uint32_t bits = 2;
  while ((1u << bits) < n) ++bits;
  pk.half_bits = (bits + 1) >> 1;
  pk.n = n;
  return pk;
}

// Bijection on [0, n): balanced Feistel on 2h bits, cycle-walked into range.
HEA_HD uint32_t hea_perm_apply(const hea_perm_key* pk, uint32_t i) {
  const uint32_t h = pk->half_bits;
  const uint32_t mask = (1u << h) - 1u;
  uint32_t x = i;
  do {
    uint32_t l = x >> h, r = x & mask;
#pragma unroll
    for (int q = 0; q < HEA_FEISTEL_ROUNDS; ++q) {
      const uint32_t f = hea_mix32(r ^ pk->k[q]) & mask;
      const uint32_t nl = r;
      r = l ^ f;
      l = nl;
    }
    x = (l << h) | r;
  } while (x >= pk->n);
  return x;
}

// Species at position q of the sorted multiset [0]*c0 ++ [1]*c1 ++ ...
HEA_HD uint32_t hea_species_of_rank(const int32_t* prefix, int k, uint32_t q) {
  uint32_t s = 0;
  while (s + 1 < (uint32_t)k && q >= (uint32_t)prefix[s + 1]) ++s;
  return s;
}
