// Greedy best-of-T local search (the hot path), one workgroup per lattice.
//
// Semantics follow pyhea/cpp/accelerate.cpp:601-694 (local_parallel_monte_carlo)
// and :498-570 (calculate_fitness_incremental):
//   * gamma starts from a fresh recount, gamma = R_s(pair count)
//     (accelerate.cpp:636-640), while prev_fit starts from the stored fitness
//     (accelerate.cpp:632);
//   * each round draws `task` valid proposals (unlike species) from the
//     lattice's stream, evaluates each incrementally from prev_gamma, keeps
//     the first strict minimum (accelerate.cpp:663-674), accepts it iff it is
//     strictly below prev_fit (accelerate.cpp:678) and carries the winner's
//     drifted gamma forward (accelerate.cpp:682);
//   * the search ends after `depth` consecutive rejected rounds.
//
// Work layout.  A workgroup has S x TP lanes: lane (s, j) handles shell s of
// proposal j (TP = proposals per pass, a multiple of 64, so every wave holds
// one shell and walks neighbour lists of one uniform length).  A round has
// two stages:
//   1. screen: every lane walks its proposal's two neighbour lists (species
//      histograms MA, MB and adjacency counts), keeps them in LDS and forms a
//      division-free estimate x_s of the shell's change of the fitness
//      squared; X = sum_s w_s x_s.
//   2. exact: wave 0 takes every proposal whose X lies within 2*beta of the
//      smallest X (normally one) and evaluates it bit-exactly as the
//      reference does: the shell's partial sum e_s = sum_ij (alpha_sij -
//      t_sij)^2 in the reference's (i, j) order, err = ((0 + w0 e0) + w1 e1)
//      + w2 e2 and F = sqrt(err) (accelerate.cpp:557-569).  The first strict
//      minimum of F over these candidates is the reference's choice, because
//      beta bounds |err - (E + X)| for every proposal (E common to all) plus a
//      margin that keeps F ties inside the window (proof beside
//      refresh_entry).
// The lattice's species sit in LDS when they fit, otherwise they are read
// from HBM.
//
// Exact FP64 bookkeeping.  The reference updates gamma by +-delta_s with
// delta_s = 1/(Z_s*N) (accelerate.cpp:526-548); the result depends on the
// order of those additions.  Every update to entry (p,q) is paired with the
// same update to (q,p), so gamma stays bit-symmetric and only rows a = sp[a1]
// and b = sp[a2] change.  For one proposal the updates an entry receives are:
//   (a,x), x not in {a,b}: -delta  MA[x] times, then +delta MB[x] times
//   (b,x), x not in {a,b}: +delta  MA[x] times, then -delta MB[x] times
//   (a,a): -delta 2*MA[a] times, then +delta 2*(MB[a]-c21+c22) times
//   (b,b): +delta 2*(MA[b]-c12+c11) times, then -delta 2*MB[b] times
//   (a,b)=(b,a): interleaved in neighbour order (a2 adjacent to a1 gives -,+)
// where MA/MB count the species among a1's/a2's shell neighbours and cXY
// counts site Y in X's list.  Inside one binade of the entry's value g with
// no rounding tie, fl(w +- delta) = w +- r for the fixed step
// r = fl(g + delta) - g, so any such sequence of at most 2*Z_s steps ends at
// g + net*r exactly, computed by one fma.  Entries whose window of reachable
// values is not provably inside one binade, or whose step ties, are marked
// r = 0 and replayed addition by addition.  Either way the result equals the
// reference's doubles bit for bit.  No FP contraction is allowed in this file
// (-ffp-contract=off); the screen uses explicit fma where its error bound
// allows.

#include <cmath>
#include <type_traits>

#include "hea_internal.h"
#include "hea_screen.h"

namespace {

constexpr int kMaxWaves = HEA_MC_MAX_THREADS / 64;  // 16
constexpr int kMapWords = 1024;  // 32768-bit conflict map (guided mode)

// Bit of site i in the conflict map (Fibonacci hash; collisions only make
// the map more conservative).
__device__ __forceinline__ uint32_t map_bit(int i) { return ((uint32_t)i * 0x9e3779b1u) >> 17; }

struct McShared {  // wcount and wmin are read as 16-byte vectors
  int wcount[kMaxWaves];
  double wmin[kMaxWaves];  // per-wave smallest screen value
  unsigned long long bmax[HEA_MAX_S];   // bits of the largest per-entry screen error bound
  unsigned long long ubmax[HEA_MAX_S];  // bits of the largest bound on |d'| (see refresh_entry)
  double beta;                          // candidate window half-width of this round
  unsigned long long res_key;           // exact stage: winner's fitness bits
  int res_j, res_a, res_b;
  int res_kind;                         // 0 reject, 1 exact accept, 2 accept decided by the window
  // prev_fit and its state (wave 0 decides; see the exact stage):
  // pfm 0: the stored fitness, not necessarily F of the recounted gamma;
  // 1: F(gamma); 2: F(gamma), not yet computed.  fest ~ err of gamma.
  double pf, fest;
  int pfm;
  uint32_t next_pos;
  int ncl;        // entries of the candidate list (a superset, see the combine step)
  unsigned gmax;  // guided mode: float bits of max_s Z_s |H_s| (gain scale)
};

// Species histograms of the shell-s neighbour lists of a1 and a2 (MA, MB:
// += 1 << (W * species)) and the adjacency counts cXY (occurrences of site Y
// in X's list).  Lists are padded to a multiple of 4 entries with -1.  NC =
// 16-byte chunks of both lists; chunks C0..C0+NG-1 are loaded before any is
// used, so their loads are in flight together.  Entries at or beyond z are
// masked: a whole list ends inside its last chunk (padding -1, which matches
// no site); a part of a split list (PART) may end anywhere, and its loads
// past the end read in-bounds entries of the next shell or row.
template <int KMAX, bool SELF, bool PART, int C0, int NG, bool LAST>
__device__ __forceinline__ void walk_group(const int32_t* __restrict__ r1,
                                           const int32_t* __restrict__ r2, int z,
                                           const uint8_t* __restrict__ sp, int a1, int a2,
                                           typename Pack<KMAX>::T& ma, typename Pack<KMAX>::T& mb,
                                           int& x11, int& x12, int& x21, int& x22) {
  using PK = Pack<KMAX>;
  int4 u[NG], v[NG];
#pragma unroll
  for (int c = 0; c < NG; ++c) {
    u[c] = reinterpret_cast<const int4*>(r1)[C0 + c];
    v[c] = reinterpret_cast<const int4*>(r2)[C0 + c];
  }
#pragma unroll
  for (int c = 0; c < NG; ++c) {
    const int nu[4] = {u[c].x, u[c].y, u[c].z, u[c].w};
    const int nv[4] = {v[c].x, v[c].y, v[c].z, v[c].w};
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const bool ok = PART ? (4 * (C0 + c) + t < z)
                           : (!(LAST && c == NG - 1) || (4 * (C0 + c) + t < z));
      const typename PK::T fa = PK::one((int)sp[ok ? nu[t] : a1]);
      const typename PK::T fb = PK::one((int)sp[ok ? nv[t] : a2]);
      ma += ok ? fa : (typename PK::T)0;
      mb += ok ? fb : (typename PK::T)0;
      x12 += (!PART || ok) && (nu[t] == a2);
      x21 += (!PART || ok) && (nv[t] == a1);
      if (SELF) {
        x11 += (!PART || ok) && (nu[t] == a1);
        x22 += (!PART || ok) && (nv[t] == a2);
      }
    }
    // one chunk's LDS reads at a time (the group's row loads are in flight)
    __builtin_amdgcn_sched_barrier(0);
  }
}

// NC chunks of both lists (z entries valid) in groups of at most 4 chunks
// (registers bound the loads in flight).
template <int KMAX, bool SELF, bool PART, int NC>
__device__ __forceinline__ void walk_pair(const int32_t* __restrict__ r1,
                                          const int32_t* __restrict__ r2, int z,
                                          const uint8_t* __restrict__ sp, int a1, int a2,
                                          typename Pack<KMAX>::T& MA, typename Pack<KMAX>::T& MB,
                                          int& c11, int& c12, int& c21, int& c22) {
  typename Pack<KMAX>::T ma = 0, mb = 0;
  int x11 = 0, x12 = 0, x21 = 0, x22 = 0;
  if constexpr (NC <= 4) {
    walk_group<KMAX, SELF, PART, 0, NC, true>(r1, r2, z, sp, a1, a2, ma, mb, x11, x12, x21, x22);
  } else {
    walk_group<KMAX, SELF, PART, 0, 4, false>(r1, r2, z, sp, a1, a2, ma, mb, x11, x12, x21, x22);
    walk_group<KMAX, SELF, PART, 4, NC - 4, true>(r1, r2, z, sp, a1, a2, ma, mb, x11, x12, x21,
                                                  x22);
  }
  MA = ma;
  MB = mb;
  c11 = x11;
  c12 = x12;
  c21 = x21;
  c22 = x22;
}

template <int KMAX, bool SELF, bool PART>
__device__ __forceinline__ void walk_pair(const int32_t* __restrict__ r1,
                                          const int32_t* __restrict__ r2, int z, int nch,
                                          const uint8_t* __restrict__ sp, int a1, int a2,
                                          typename Pack<KMAX>::T& MA, typename Pack<KMAX>::T& MB,
                                          int& c11, int& c12, int& c21, int& c22) {
  switch (nch) {  // uniform across the wave (one shell and part size per wave)
#define HEA_WALK(n) \
  case n: walk_pair<KMAX, SELF, PART, n>(r1, r2, z, sp, a1, a2, MA, MB, c11, c12, c21, c22); break;
    HEA_WALK(1) HEA_WALK(2) HEA_WALK(3) HEA_WALK(4) HEA_WALK(5) HEA_WALK(6) HEA_WALK(7) HEA_WALK(8)
#undef HEA_WALK
    default: break;
  }
}

// Occurrences of site a2 among the z entries of list r1 (NC 16-byte chunks,
// all loads in flight together); entries at or beyond z are masked for a
// part of a split list (a whole list ends in its padding, -1).
template <bool PART, int NC>
__device__ __forceinline__ int count_site(const int32_t* __restrict__ r1, int z, int a2) {
  int4 u[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) u[c] = reinterpret_cast<const int4*>(r1)[c];
  int n = 0;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const int nu[4] = {u[c].x, u[c].y, u[c].z, u[c].w};
#pragma unroll
    for (int t = 0; t < 4; ++t) n += ((!PART || 4 * c + t < z) && nu[t] == a2) ? 1 : 0;
  }
  return n;
}

template <bool PART>
__device__ __forceinline__ int count_site(const int32_t* __restrict__ r1, int z, int nch, int a2) {
  switch (nch) {  // uniform across the wave
#define HEA_COUNT(n) \
  case n: return count_site<PART, n>(r1, z, a2);
    HEA_COUNT(1) HEA_COUNT(2) HEA_COUNT(3) HEA_COUNT(4) HEA_COUNT(5) HEA_COUNT(6) HEA_COUNT(7)
    HEA_COUNT(8)
#undef HEA_COUNT
    default: return 0;
  }
}

// Shell-s species histogram of a neighbour list in the walk's packing, from
// the list owner's neighbour record r (HeaHistLayout: the shell's fields
// start at bit o, are w bits wide and W bits apart); z = Z_s.
template <int KMAX>
__device__ __forceinline__ typename Pack<KMAX>::T rec_pack(unsigned long long r, int o, int w,
                                                           int W, int K, int z) {
  using PK = Pack<KMAX>;
  typename PK::T m = 0;
  int rest = z;
  const unsigned long long fm = (1ull << w) - 1ull;
#pragma unroll
  for (int x = 0; x < KMAX - 1; ++x) {
    if (x < K - 1) {
      const int f = (int)((r >> (x * W + o)) & fm);
      rest -= f;
      m += (typename PK::T)f << (PK::W * x);
    }
  }
  return m + ((typename PK::T)rest << (PK::W * (K - 1)));
}

// LDS carve-up of one workgroup (byte offsets; every region 16-byte aligned).
template <int KMAX>
struct LdsLayout {
  static constexpr int TAB = HEA_MAX_S * KMAX * KMAX;
  using PT = typename Pack<KMAX>::T;
  size_t G, E0, R, tgt, SC, H, conc, xs, ex, xw, prop, gkey, cmap, WA, WB, WC, cl, sl, sh, sp, total;
  __host__ __device__ LdsLayout(int task, int n, bool sp_lds) {
    size_t o = 0;
    auto take = [&o](size_t bytes) {
      const size_t r = o;
      o = (o + bytes + 15) & ~(size_t)15;
      return r;
    };
    G = take(sizeof(double) * TAB);
    E0 = take(sizeof(double) * TAB);
    R = take(sizeof(double) * TAB);
    tgt = take(sizeof(double) * TAB);
    SC = take(sizeof(double2) * TAB);
    H = take(sizeof(float) * HEA_MAX_S * KMAX * 8);
    conc = take(sizeof(double) * KMAX);
    xs = take(sizeof(double) * HEA_MAX_S * (size_t)task);
    ex = take(sizeof(double) * 64);
    xw = take(sizeof(double) * 64 * (4 + KMAX));
    prop = take(sizeof(int2) * (size_t)task);
    gkey = take(sizeof(unsigned long long) * (size_t)task);
    cmap = take(sizeof(uint32_t) * kMapWords);
    WA = take(sizeof(PT) * HEA_MAX_S * (size_t)task);
    WB = take(sizeof(PT) * HEA_MAX_S * (size_t)task);
    WC = take(sizeof(uint32_t) * HEA_MAX_S * (size_t)task);
    cl = take(sizeof(int) * (size_t)task);
    sl = take(sizeof(int) * 64);
    sh = take(sizeof(McShared));
    sp = o;  // species [N] when resident, else the proposals' species pairs [task]
    total = (o + (sp_lds ? (size_t)n : sizeof(uint16_t) * (size_t)task) + 15) & ~(size_t)15;
  }
};

template <int KMAX>
struct Lds {
  static constexpr int TAB = LdsLayout<KMAX>::TAB;
  using PT = typename Pack<KMAX>::T;
  double* G;     // [S][KMAX][KMAX] prev_gamma
  double* E0;    // [S][KMAX][KMAX] (1 - G/d - t)^2 of prev_gamma
  double* R;     // [S][KMAX][KMAX] safe step or 0
  double* tgt;   // [S][KMAX][KMAX] target alpha
  double2* SC;   // [S][KMAX][KMAX] {d_ij + d_ji, -delta_s/dd_ij}, d = alpha - t
  float* H;      // [S][KMAX][8] guided mode: dF2/dn + its transpose (rows padded to 8)
  double* conc;  // [KMAX] fl(count/N)
  double* xs;    // [S][task] screen values; row 0 then holds the combined X
  double* ex;    // [64] exact-stage shell sums (wave 0)
  double* xw;    // [64][4 + KMAX] exact stage, lane (c, s, x): gamma'(a,x), gamma'(b,x),
                 // their alphas, then the terms of row x
  int2* prop;    // [task]
  unsigned long long* gkey;  // [task] guided mode: best (score, candidate) of each slot
  uint32_t* cmap;  // [kMapWords] guided mode: hashed sites touched by this round's accepts
  PT* WA;        // [S][task] walk results: histogram of a1's neighbours
  PT* WB;        // [S][task] histogram of a2's neighbours
  uint32_t* WC;  // [S][task] walk_pack(c11, c12, c21, c22, a, b)
  int* cl;       // [task] exact-stage candidate list
  int* sl;       // [64] guided mode: the round's shortlist for further accepts
  McShared* sh;
  uint8_t* sp;   // [N] when resident
  uint16_t* pab; // [task] species in HBM: a | b << 8 of each proposal, from the draw
  static size_t bytes(int task, int n, bool sp_lds) { return LdsLayout<KMAX>(task, n, sp_lds).total; }
  __device__ void carve(unsigned char* base, int task, bool sp_lds) {
    const LdsLayout<KMAX> L(task, 0, sp_lds);
    G = reinterpret_cast<double*>(base + L.G);
    E0 = reinterpret_cast<double*>(base + L.E0);
    R = reinterpret_cast<double*>(base + L.R);
    tgt = reinterpret_cast<double*>(base + L.tgt);
    SC = reinterpret_cast<double2*>(base + L.SC);
    H = reinterpret_cast<float*>(base + L.H);
    conc = reinterpret_cast<double*>(base + L.conc);
    xs = reinterpret_cast<double*>(base + L.xs);
    ex = reinterpret_cast<double*>(base + L.ex);
    xw = reinterpret_cast<double*>(base + L.xw);
    prop = reinterpret_cast<int2*>(base + L.prop);
    gkey = reinterpret_cast<unsigned long long*>(base + L.gkey);
    cmap = reinterpret_cast<uint32_t*>(base + L.cmap);
    WA = reinterpret_cast<PT*>(base + L.WA);
    WB = reinterpret_cast<PT*>(base + L.WB);
    WC = reinterpret_cast<uint32_t*>(base + L.WC);
    cl = reinterpret_cast<int*>(base + L.cl);
    sl = reinterpret_cast<int*>(base + L.sl);
    sh = reinterpret_cast<McShared*>(base + L.sh);
    sp = sp_lds ? base + L.sp : nullptr;
    pab = sp_lds ? nullptr : reinterpret_cast<uint16_t*>(base + L.sp);
  }
  __device__ __forceinline__ int idx(int s, int i, int j) const { return (s * KMAX + i) * KMAX + j; }
  __device__ __forceinline__ int hidx(int s, int i, int j) const { return (s * KMAX + i) * 8 + j; }
};

template <int KMAX>
__device__ __forceinline__ bool touched(const Lds<KMAX>& T, int2 pr) {
  const uint32_t b1 = map_bit(pr.x), b2 = map_bit(pr.y);
  return ((T.cmap[b1 >> 5] >> (b1 & 31u)) | (T.cmap[b2 >> 5] >> (b2 & 31u))) & 1u;
}

// Orders LDS traffic between the lanes of one wave (the hardware keeps a
// wave's LDS operations in order; this stops the compiler from moving them).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Sum of v over the 2^lg adjacent lanes that share a proposal (lg <= 2, the
// same for the whole wave): DPP quad permutations, every source in the quad.
__device__ __forceinline__ uint32_t part_sum(uint32_t v, int lg) {
  if (lg > 0) v += (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0xB1, 0xf, 0xf, false);
  if (lg > 1) v += (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x4E, 0xf, 0xf, false);
  return v;
}
// Packed histograms: fields never exceed HEA_MAX_Z, so sums carry into no
// neighbouring field and the two halves of a 64-bit pack add separately.
__device__ __forceinline__ uint64_t part_sum(uint64_t v, int lg) {
  return (uint64_t)part_sum((uint32_t)v, lg) | ((uint64_t)part_sum((uint32_t)(v >> 32), lg) << 32);
}

// Walk record of one (proposal, shell): the adjacency counts c11, c12, c21,
// c22 (each <= HEA_MAX_Z, 5 bits) and the species a, b of the two sites.
struct WalkRec {
  int c11, c12, c21, c22, a, b;
};
__device__ __forceinline__ uint32_t walk_pack(int c11, int c12, int c21, int c22, int a, int b) {
  return (uint32_t)c11 | ((uint32_t)c12 << 5) | ((uint32_t)c21 << 10) | ((uint32_t)c22 << 15) |
         ((uint32_t)a << 20) | ((uint32_t)b << 24);
}
__device__ __forceinline__ WalkRec walk_unpack(uint32_t w) {
  return WalkRec{(int)(w & 31u),         (int)((w >> 5) & 31u), (int)((w >> 10) & 31u),
                 (int)((w >> 15) & 31u), (int)((w >> 20) & 15u), (int)((w >> 24) & 15u)};
}

// One step of a DPP reduction: v = min(v, v of the lane DPP control CTRL
// names), in the rows ROWS enables; lanes of other rows keep v.  Every
// control used here names a lane of the same wave, so no lane reads an
// out-of-range source.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_min_step(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const int lo = (int)(uint32_t)b, hi = (int)(uint32_t)(b >> 32);
  const int olo = __builtin_amdgcn_update_dpp(lo, lo, CTRL, ROWS, 0xf, false);
  const int ohi = __builtin_amdgcn_update_dpp(hi, hi, CTRL, ROWS, 0xf, false);
  const double o = __longlong_as_double(
      (long long)(((unsigned long long)(uint32_t)ohi << 32) | (uint32_t)olo));
  return fmin(v, o);
}

// Smallest v over all 64 lanes (every lane active), returned to every lane.
// row_ror:1,2,4,8 leave each 16-lane row's minimum in all its lanes;
// row_bcast:15 folds rows 0 and 2 into rows 1 and 3, row_bcast:31 folds
// row 1 into rows 2 and 3, so lane 63 holds the wave's minimum.
__device__ __forceinline__ double wave_min(double v) {
  v = dpp_min_step<0x121, 0xf>(v);
  v = dpp_min_step<0x122, 0xf>(v);
  v = dpp_min_step<0x124, 0xf>(v);
  v = dpp_min_step<0x128, 0xf>(v);
  v = dpp_min_step<0x142, 0xa>(v);
  v = dpp_min_step<0x143, 0xc>(v);
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, 63);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), 63);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Largest v over all 64 lanes (every lane active), returned to every lane.
__device__ __forceinline__ double wave_max(double v) { return -wave_min(-v); }

// Refreshes the tables of entry (s,i,j) (refresh_entry_core) and raises the
// workgroup's bound maxima by its bounds.
template <int KMAX, bool GUIDED>
__device__ void refresh_entry(const HeaMcArgs& A, const Lds<KMAX>& T, int s, int i, int j) {
  const EntryBounds r = refresh_entry_core<KMAX, GUIDED>(A, T, s, i, j);
  atomicMax(&T.sh->bmax[s], (unsigned long long)__double_as_longlong(r.P));
  atomicMax(&T.sh->ubmax[s], (unsigned long long)__double_as_longlong(r.ub));
  if constexpr (GUIDED) atomicMax(&T.sh->gmax, __float_as_uint(r.hz));
}

// Guided mode: candidate item `ctr` of a lattice's candidate stream (key
// gk), in the proposal slot whose first item is `sctr`: two sites (s1, s2)
// of n, drawn by multiply-high from independent hashes, and a 32-bit word h3
// for the Gumbel noise.  blk == 0: both sites uniform over the lattice.
// blk > 0: the slot draws two blocks of 2^blk consecutive sites once (from
// sctr; the last block of the lattice may be shorter) and each candidate one
// site in each block, so a slot's candidates read their neighbour records
// from two spans of 8 << blk bytes.  This stream feeds only the guided
// proposal choice; the exact acceptance never depends on it.
__device__ __forceinline__ void guided_pair(uint32_t gk, uint32_t ctr, uint32_t sctr, uint32_t n,
                                            int blk, uint32_t* s1, uint32_t* s2, uint32_t* h3) {
  const uint32_t h0 = hea_mix32(gk ^ (ctr * 0x9e3779b9u));
  const uint32_t h1 = hea_mix32(h0 ^ 0x85ebca6bu);
  const uint32_t h2 = hea_mix32(h1 ^ 0xc2b2ae35u);
  if (blk == 0) {
    *s1 = __umulhi(h0, n);
    *s2 = __umulhi(h1, n);
  } else {
    const uint32_t k0 = hea_mix32(gk ^ (sctr * 0x9e3779b9u) ^ 0x27d4eb2fu);
    const uint32_t k1 = hea_mix32(k0 ^ 0x165667b1u);
    const uint32_t bs = 1u << blk, nb = (n + bs - 1) >> blk;
    const uint32_t x = __umulhi(k0, nb) << blk, y = __umulhi(k1, nb) << blk;
    *s1 = x + __umulhi(h0, min(n - x, bs));
    *s2 = y + __umulhi(h1, min(n - y, bs));
  }
  *h3 = h2;
}

// Fitness of the current gamma as the reference computes it
// (accelerate.cpp:557-569): E0 holds the current terms; lane s < S adds its
// shell's K*K terms in row-major order.  Called by all lanes of one wave.
template <int KMAX, bool EXACT>
__device__ __forceinline__ double current_fitness(const HeaMcArgs& A, const Lds<KMAX>& T,
                                                  int lane) {
  const int K = A.K, S = A.S;
  if (lane < S) {
    const double* e0 = T.E0 + T.idx(lane, 0, 0);
    double es = 0.0;
#pragma unroll
    for (int i = 0; i < KMAX; ++i)
#pragma unroll
      for (int y = 0; y < KMAX; ++y)
        if (EXACT || (i < K && y < K)) es = es + e0[i * KMAX + y];
    T.ex[lane] = es;
  }
  wave_sync();
  double err = 0.0;
  for (int t = 0; t < S; ++t) err = err + A.weight[t] * T.ex[t];
  wave_sync();  // ex is reused
  return sqrt(err);
}

__device__ __forceinline__ bool key_less(unsigned long long f1, int j1, unsigned long long f2,
                                         int j2) {
  return (f1 < f2) || (f1 == f2 && j1 < j2);
}

// Exact stage of a round, run by one wave (wave 0): the candidates T.cl[0,
// nsup) (a superset of the proposals whose X lies within 2 beta of xmin),
// first strict minimum of F (accelerate.cpp:663-674) and the accept test
// (:678).  Writes the decision to T.sh (res_kind, res_j, res_a, res_b,
// res_key, pf, pfm, fest) and, for an accept, the winner's new gamma rows;
// with `reset` an accept also clears the bound maxima that the refresh of
// every entry then takes again.
template <int KMAX, bool EXACT>
__device__ __forceinline__ void exact_stage(const HeaMcArgs& A, const Lds<KMAX>& T,
                                            const uint8_t* __restrict__ sp, int lane, int task,
                                            int nsup, double xmin, double beta, bool reset) {
  const int K = A.K, S = A.S;
  double prev_fit = T.sh->pf, fest = T.sh->fest;
  int pfm = T.sh->pfm;
  double thr = xmin + 2.0 * beta;
  if (!(thr < __builtin_inf())) thr = __builtin_inf();
  // filter the waves' lists against the workgroup's window, compacting
  // in place (every entry is read before any of its chunk is written)
  int nc = 0, jfirst = -1;  // jfirst = cl[0] after the compaction
  for (int k0 = 0; k0 < nsup; k0 += 64) {
    const int j = (k0 + lane < nsup) ? T.cl[k0 + lane] : -1;
    const bool f = j >= 0 && !(T.xs[j] > thr);
    const unsigned long long bal = __ballot(f);
    if (nc == 0 && bal) jfirst = __builtin_amdgcn_readlane(j, (int)__builtin_ctzll(bal));
    wave_sync();
    if (f) T.cl[nc + __popcll(bal & ((1ull << lane) - 1ull))] = j;
    nc += __popcll(bal);
  }
  wave_sync();
  // lane (c, s, x): candidate c of the chunk, shell s, species x
  const int GS = S * K;      // lanes per candidate
  const int cpc = 64 / GS;   // candidates per chunk
  const int g = lane / K, x = lane - g * K;
  const int c = g / S, s = g - c * S;
  constexpr int XW = 4 + KMAX;
  double* slot = T.xw + lane * XW;
  const double* grp = T.xw + g * K * XW;  // lane (c, s, y) at grp + y * XW
  unsigned long long bk = ~0ull;
  int bj = 0x7fffffff, bc = 0, wc0 = -1;
  int ra = 0, rb = 0;  // window-decided accept: the winner's species (lanes < S*K)
  // When prev_fit = F(gamma), the window decides most rounds: every
  // proposal's err lies within beta of E + X (E common, X = 0 for the
  // current lattice), and 2 beta keeps sqrt's rounding out of the way.
  // One candidate with X < -2 beta is certainly accepted; X > 2 beta for
  // the smallest X is certainly a reject.  Near the threshold, or
  // otherwise, the exact path decides.
  const double mg = 2.0 * beta;
  const bool near = A.thr > 0.0 && !(fest + xmin > A.thr * A.thr * 1.01);
  int kind = -1;
  if (pfm != 0 && !near) {
    if (nc == 1 && xmin + mg < 0.0) kind = 2;
    else if (xmin - mg > 0.0) kind = 0;
  }
  if (kind == -1 && pfm == 2) {  // the exact path compares with prev_fit
    prev_fit = current_fitness<KMAX, EXACT>(A, T, lane);
    pfm = 1;
  }
  if (kind == 0) {
    if (lane == 0) T.sh->res_kind = 0;
    return;
  }
  for (int c0 = 0; c0 < nc; c0 += cpc) {
    int j = -1;
    if (nc == 1)  // the usual case: the candidate came with the filter's ballot
      j = (lane < GS) ? jfirst : -1;
    else if (lane < cpc * GS && c0 + c < nc)
      j = T.cl[c0 + c];
    int a = 0, b = 0;
    double nA = 0.0, nB = 0.0;
    if (j >= 0) {
      const int2 pr = T.prop[j];  // only the rare replay path reads the sites
      const int w = s * task + j;
      const WalkRec r = walk_unpack(T.WC[w]);
      a = r.a;
      b = r.b;
      exact_entries<KMAX>(A, T, sp, s, pr.x, pr.y, a, b, x, T.WA[w], T.WB[w], r.c11, r.c12,
                          r.c21, r.c22, nA, nB);
      slot[0] = nA;
      slot[1] = nB;
      if (kind != 2) {  // alphas feed the terms, which a decided accept skips
        slot[2] = 1.0 - nA / (T.conc[a] * T.conc[x]);
        slot[3] = 1.0 - nB / (T.conc[b] * T.conc[x]);
      }
    }
    wave_sync();
    if (kind == 2) {  // the one candidate is accepted: its new rows only
      if (j >= 0) {
        T.G[T.idx(s, a, x)] = nA;
        T.G[T.idx(s, x, a)] = nA;
        T.G[T.idx(s, b, x)] = nB;
        T.G[T.idx(s, x, b)] = nB;
      }
      bj = jfirst;
      ra = a;
      rb = b;
      break;
    }
    if (j >= 0) {  // terms of row x (accelerate.cpp:560-566)
      const double* tg = T.tgt + T.idx(s, x, 0);
      const double* e0 = T.E0 + T.idx(s, x, 0);
      const double sAx = slot[2], sBx = slot[3];
#pragma unroll
      for (int y = 0; y < KMAX; ++y) {
        if (EXACT || y < K) {
          const bool chg = (x == a) || (x == b) || (y == a) || (y == b);
          const double sro = (x == a)   ? grp[y * XW + 2]
                             : (x == b) ? grp[y * XW + 3]
                             : (y == a) ? sAx
                                        : sBx;
          const double df = sro - tg[y];
          slot[4 + y] = chg ? df * df : e0[y];
        }
      }
    }
    wave_sync();
    if (j >= 0 && x == 0) {  // the shell's K*K terms in row-major order
      double es = 0.0;
#pragma unroll
      for (int i = 0; i < KMAX; ++i)
#pragma unroll
        for (int y = 0; y < KMAX; ++y)
          if (EXACT || (i < K && y < K)) es = es + grp[i * XW + 4 + y];
      T.ex[g] = es;
    }
    wave_sync();
    unsigned long long rk = ~0ull;
    int rj = 0x7fffffff, rc = 0;
    if (j >= 0 && s == 0 && x == 0) {
      double err = 0.0;
      for (int t = 0; t < S; ++t) err = err + A.weight[t] * T.ex[c * S + t];
      rk = (unsigned long long)__double_as_longlong(sqrt(err));
      rj = j;
      rc = c;
    }
    if (nc - c0 == 1) {  // the usual case: one candidate, in lane 0
      rk = ((unsigned long long)(uint32_t)__builtin_amdgcn_readfirstlane((int)(rk >> 32)) << 32) |
           (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)rk);
      rj = __builtin_amdgcn_readfirstlane(rj);
      rc = 0;
    } else {
#pragma unroll
      for (int m = 32; m >= 1; m >>= 1) {
        const unsigned long long ok =
            ((unsigned long long)(uint32_t)__shfl_xor((int)(rk >> 32), m) << 32) |
            (uint32_t)__shfl_xor((int)(uint32_t)rk, m);
        const int oj = __shfl_xor(rj, m);
        const int oc = __shfl_xor(rc, m);
        if (key_less(ok, oj, rk, rj)) {
          rk = ok;
          rj = oj;
          rc = oc;
        }
      }
    }
    if (key_less(rk, rj, bk, bj)) {
      bk = rk;
      bj = rj;
      bc = rc;
      wc0 = c0;
    }
    if (c0 + cpc < nc) wave_sync();  // the next chunk reuses xw and ex
  }
  if (kind == 2) {
    if (reset && lane < HEA_MAX_S) {  // the refresh below takes new maxima
      T.sh->bmax[lane] = 0ull;
      T.sh->ubmax[lane] = 0ull;
    }
    if (lane == 0) {
      if (reset) T.sh->gmax = 0u;
      T.sh->res_kind = 2;
      T.sh->res_j = bj;
      T.sh->res_a = ra;
      T.sh->res_b = rb;
      T.sh->pfm = 2;
      T.sh->fest = fest + xmin;
    }
    return;
  }
  const double best_fit = __longlong_as_double((long long)bk);
  int a = 0, b = 0;
  if (best_fit < prev_fit) {  // accept: write the winner's rows
    const int2 pr = T.prop[bj];
    {
      const WalkRec r = walk_unpack(T.WC[bj]);  // shell 0's walk holds the species
      a = r.a;
      b = r.b;
    }
    if (wc0 + cpc < nc) {  // a later chunk overwrote the winner's entries
      if (lane < GS) {
        const int w = s * task + bj;
        const WalkRec r = walk_unpack(T.WC[w]);
        double nA, nB;
        exact_entries<KMAX>(A, T, sp, s, pr.x, pr.y, a, b, x, T.WA[w], T.WB[w], r.c11,
                            r.c12, r.c21, r.c22, nA, nB);
        slot[0] = nA;
        slot[1] = nB;
      }
      bc = 0;
    }
    wave_sync();
    // rows a and b of every shell, with their mirrors (columns a and b)
    for (int e = lane; e < S * 2 * K; e += 64) {
      const int t = e / (2 * K), h = (e / K) & 1, y = e % K;
      const double v = T.xw[((bc * S + t) * K + y) * XW + h];
      const int r = h ? b : a;
      T.G[T.idx(t, r, y)] = v;
      T.G[T.idx(t, y, r)] = v;
    }
    if (reset && lane < HEA_MAX_S) {  // the refresh below takes new maxima
      T.sh->bmax[lane] = 0ull;
      T.sh->ubmax[lane] = 0ull;
    }
    if (reset && lane == 0) T.sh->gmax = 0u;
  }
  if (lane == 0) {
    const bool acc1 = best_fit < prev_fit;
    T.sh->res_kind = acc1 ? 1 : 0;
    T.sh->res_key = bk;
    T.sh->res_j = bj;
    T.sh->res_a = a;
    T.sh->res_b = b;
    T.sh->pf = acc1 ? best_fit : prev_fit;
    T.sh->pfm = acc1 ? 1 : pfm;
    if (acc1) T.sh->fest = best_fit * best_fit;
  }
}

// Guided mode: first-order change of F^2 when site i (species a, neighbour
// record ri) and site j (species b, record rj) swap species:
// sum_s sum_d (H_s[b][d] - H_s[a][d]) (cnt_s[i][d] - cnt_s[j][d]) = u[i][b] +
// u[j][a] in the site gains u of the relaxation's gradient.  Absent shells
// and species have zero rows in H; the last species' count difference is
// minus the others' (both lists hold Z_s).
template <int KMAX>
__device__ __forceinline__ float guided_gain(const Lds<KMAX>& T, const HeaHistLayout& hl, int K,
                                             int a, int b, unsigned long long ri,
                                             unsigned long long rj) {
  float d = 0.0f;
#pragma unroll
  for (int sh = 0; sh < HEA_MAX_S; ++sh) {
    const float* Ha = T.H + T.hidx(sh, a, 0);
    const float* Hb = T.H + T.hidx(sh, b, 0);
    const unsigned long long fm = (1ull << hl.w[sh]) - 1ull;
    int rest = 0;
#pragma unroll
    for (int x = 0; x < KMAX - 1; ++x) {
      const int sft = min(x * hl.W + hl.o[sh], 63);  // in range; x >= K-1 is masked
      const int dc = (x < K - 1) ? (int)((ri >> sft) & fm) - (int)((rj >> sft) & fm) : 0;
      rest -= dc;
      d += (Hb[x] - Ha[x]) * (float)dc;
    }
    const float* Hl = Hb + (K - 1);  // species K-1
    d += (*Hl - Ha[K - 1]) * (float)rest;
  }
  return d;
}

// Guided mode: the neighbours of a swap's two sites, work item t of
// [0, 2 sum Z) (a1's lists, then a2's) in lane t mod stride, slot t / stride;
// -1 where there is no item.  2 sum Z <= 186 <= 3 x 64, so three slots cover
// every item for any stride >= 64.
struct SwapNbrs {
  int nb[3];
};

// Shell of item q of a site's concatenated lists (q < sum Z; absent shells
// have Z = 0), and q made the index within that shell.  Compares and
// selects of the shell sizes: indexing them by a lane-varying shell would
// load from memory.
__device__ __forceinline__ int item_shell(const HeaMcArgs& A, int& q) {
  const int z0 = A.Z[0], z01 = z0 + A.Z[1];
  const int s = (q >= z0) + (q >= z01);
  q -= s == 0 ? 0 : s == 1 ? z0 : z01;
  return s;
}
__device__ __forceinline__ int shell_off(const HeaMcArgs& A, int s) {
  return s == 0 ? A.off[0] : s == 1 ? A.off[1] : A.off[2];
}

__device__ __forceinline__ SwapNbrs swap_nbrs_load(const HeaMcArgs& A, int2 pr, int first,
                                                   int stride) {
  int zs = 0;
  for (int s = 0; s < A.S; ++s) zs += A.Z[s];
  SwapNbrs r;
#pragma unroll
  for (int u = 0; u < 3; ++u) {
    const int t = first + u * stride;
    r.nb[u] = -1;
    if (t < 2 * zs) {
      const int site = t < zs ? pr.x : pr.y;
      int q = t < zs ? t : t - zs;
      const int s = item_shell(A, q);
      r.nb[u] = A.nbr[(size_t)site * A.stride + shell_off(A, s) + q];
    }
  }
  return r;
}

// swap_nbrs_load from copies of the two sites' rows (a1's, then a2's, each
// A.stride entries) kept in LDS.
__device__ __forceinline__ SwapNbrs swap_nbrs_rows(const HeaMcArgs& A, const int32_t* rows,
                                                   int first, int stride) {
  int zs = 0;
  for (int s = 0; s < A.S; ++s) zs += A.Z[s];
  SwapNbrs r;
#pragma unroll
  for (int u = 0; u < 3; ++u) {
    const int t = first + u * stride;
    r.nb[u] = -1;
    if (t < 2 * zs) {
      int q = t < zs ? t : t - zs;
      const int s = item_shell(A, q);
      r.nb[u] = rows[(t < zs ? 0 : A.stride) + shell_off(A, s) + q];
    }
  }
  return r;
}

// After swapping a1 = pr.x (species a -> b) and a2 = pr.y (b -> a): every
// neighbour of the two sites (n, from swap_nbrs_load with the same first and
// stride) sees the change in its neighbour-count record (the relation is
// symmetric, so these are exactly the records that change); with MAP
// (guided mode) the two sites and their neighbours are also marked in the
// round's conflict map.
template <int KMAX, bool MAP = true>
__device__ __forceinline__ void guided_apply_swap(const HeaMcArgs& A, const Lds<KMAX>& T,
                                                  unsigned long long* hist,
                                                  const HeaHistLayout& hl, int2 pr, int a, int b,
                                                  int first, int stride, const SwapNbrs& n) {
  const int K = A.K;
  int zs = 0;
  for (int s = 0; s < A.S; ++s) zs += A.Z[s];
#pragma unroll
  for (int u = 0; u < 3; ++u) {
    if (n.nb[u] < 0) continue;
    const int t = first + u * stride;
    int q = t < zs ? t : t - zs;
    const int s = item_shell(A, q);
    // the last species has no field (its count is implied); the field is
    // selected without a variable index (which would put the layout in
    // scratch memory)
    const int os = s == 0 ? hl.o[0] : s == 1 ? hl.o[1] : hl.o[2];
    const unsigned long long fa = (a < K - 1) ? 1ull << (a * hl.W + os) : 0ull;
    const unsigned long long fb = (b < K - 1) ? 1ull << (b * hl.W + os) : 0ull;
    const unsigned long long dv = t < zs ? fb - fa : fa - fb;  // modulo 2^64
    // every access to a lattice's records comes from its own workgroup:
    // workgroup-scope atomics, loads and barriers order them
    __hip_atomic_fetch_add(hist + n.nb[u], dv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if constexpr (MAP) {
      const uint32_t mb = map_bit(n.nb[u]);
      atomicOr(&T.cmap[mb >> 5], 1u << (mb & 31u));
    }
  }
  if (MAP && first < 2) {
    const uint32_t mb = map_bit(first == 0 ? pr.x : pr.y);
    atomicOr(&T.cmap[mb >> 5], 1u << (mb & 31u));
  }
}

// After an accepted swap of species a and b: refreshes the entries that
// changed, rows and columns a and b of every shell (4K - 4 per shell), by the
// 64 lanes of one wave.  The bound maxima only grow: every other entry's
// bounds are unchanged, so the maxima stay upper bounds of the current
// entries' bounds (beta stays valid, at worst wider than needed).
template <int KMAX>
__device__ __forceinline__ void refresh_rows(const HeaMcArgs& A, const Lds<KMAX>& T, int a, int b,
                                             int lane) {
  const int K = A.K, S = A.S, per = 4 * K - 4;
  for (int e0 = 0; e0 < S * per; e0 += 64) {
    const int e = e0 + lane;
    EntryBounds r{0.0, 0.0, 0.0f};
    int s = HEA_MAX_S;
    if (e < S * per) {
      s = e / per;
      const int k = e - s * per;
      int i, j;
      if (k < 2 * K) {  // rows a and b
        i = k < K ? a : b;
        j = k < K ? k : k - K;
      } else {          // columns a and b of the other rows
        const int m = k - 2 * K, h = m / (K - 2);
        int y = m - h * (K - 2);
        y += (y >= min(a, b));
        y += (y >= max(a, b));
        i = y;
        j = h ? b : a;
      }
      r = refresh_entry_core<KMAX, true>(A, T, s, i, j);
    }
    for (int t = 0; t < S; ++t) {
      const double bm = wave_max(s == t ? r.P : 0.0);
      const double um = wave_max(s == t ? r.ub : 0.0);
      if (lane == 0) {  // non-negative doubles order like their bit patterns
        const unsigned long long nb = (unsigned long long)__double_as_longlong(bm);
        const unsigned long long nu = (unsigned long long)__double_as_longlong(um);
        if (nb > T.sh->bmax[t]) T.sh->bmax[t] = nb;
        if (nu > T.sh->ubmax[t]) T.sh->ubmax[t] = nu;
      }
    }
    const unsigned hm = __float_as_uint((float)wave_max((double)r.hz));
    if (lane == 0 && hm > T.sh->gmax) T.sh->gmax = hm;
  }
}

// Guided mode: the further accepts of a round, made by wave 0 alone, so an
// accept costs no workgroup barrier.  T.sl holds the round's shortlist: up to
// 64 proposals with predicted decreases (X < 0), the smallest X of each
// wave's share of the round's screen.  Each step drops the proposals that
// touch a site swapped in this round or one of its neighbours (their walks
// are stale), screens the others again from their stored walks against the
// refreshed tables, and sends those within 2 beta of the smallest X to the
// exact stage, which accepts the first strict minimum iff its exact fitness
// is lower than the current one (the rule of accelerate.cpp:663-689 applied
// to the remaining proposals).  Stops at the first reject, when no proposal
// is left, or after `steps` accepts.  Returns the accepts; adds the
// proposals screened to *screened; sets *hit when an accepted fitness is
// <= thr.
template <int KMAX, bool EXACT>
__device__ int guided_more_accepts(const HeaMcArgs& A, const Lds<KMAX>& T, uint8_t* sp,
                                   unsigned long long* hist, const HeaHistLayout& hl, int lane,
                                   int task, int steps, long long* screened, bool* hit) {
  const int K = A.K, S = A.S;
  int j = T.sl[lane];
  int made = 0;
  for (int step = 0; step < steps; ++step) {
    if (j >= 0 && touched(T, T.prop[j])) j = -1;  // for good: its walk is stale
    const double beta = screen_beta(A, T);
    double X = __builtin_inf();
    if (j >= 0) {
      X = 0.0;
      for (int s = 0; s < S; ++s) {
        const int w = s * task + j;
        const WalkRec r = walk_unpack(T.WC[w]);
        X = X + A.weight[s] * screen_shell<KMAX, EXACT>(T, EXACT ? KMAX : K, s, r.a, r.b,
                                                        T.WA[w], T.WB[w], r.c11, r.c12,
                                                        r.c21, r.c22);
      }
      T.xs[j] = X;
    }
    const unsigned long long live = __ballot(j >= 0);
    *screened += __popcll(live);
    const double xmin = wave_min(X);
    if (!(xmin < __builtin_inf())) break;
    const bool f = j >= 0 && !(X > xmin + 2.0 * beta);
    const unsigned long long bal = __ballot(f);
    if (f) T.cl[__popcll(bal & ((1ull << lane) - 1ull))] = j;
    wave_sync();
    exact_stage<KMAX, EXACT>(A, T, sp, lane, task, __popcll(bal), xmin, beta, false);
    wave_sync();
    const int kind = T.sh->res_kind;
    if (kind == 0) break;
    ++made;
    const int2 pr = T.prop[T.sh->res_j];
    const int a = T.sh->res_a, b = T.sh->res_b;
    if (kind == 1 && __longlong_as_double((long long)T.sh->res_key) <= A.thr) *hit = true;
    wave_sync();
    if (lane == 0) {
      sp[pr.x] = (uint8_t)b;
      sp[pr.y] = (uint8_t)a;
    }
    guided_apply_swap<KMAX>(A, T, hist, hl, pr, a, b, lane, 64, swap_nbrs_load(A, pr, lane, 64));
    refresh_rows<KMAX>(A, T, a, b, lane);
    wave_sync();
  }
  return made;
}

// Local search of lattice l of the population described by A.  GUIDED
// selects the proposal source (A.hist != nullptr): a separate instantiation,
// so the plain kernels carry none of the guided code's registers.
template <int KMAX, bool SP_LDS, bool EXACT, bool SELF, bool GUIDED>
__device__ __forceinline__ void mc_lattice(const HeaMcArgs& A, int l, unsigned char* smem) {
  using PT = typename Pack<KMAX>::T;
  const HeaConst& P = *A.pc;
  const int nthr = blockDim.x;
  const int nwaves = (nthr + 63) >> 6;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: branches on it are scalar
  const int N = A.N, K = A.K, S = A.S;
  const int task = A.task, TP = A.tp;
  // Screen lanes: shell s's group starts at lane gbase[s]; its list is walked
  // in 2^glog[s] parts by adjacent lanes sharing a proposal slot.  Groups are
  // whole waves, so every wave walks one shell with one part size.
  int s_grp = 0;
  for (int t = 1; t < S; ++t)
    if (tid >= A.gbase[t]) s_grp = t;
  const int s_me = __builtin_amdgcn_readfirstlane(s_grp);  // this wave's shell
  const bool scr = tid < A.gbase[S];                       // a screen lane (uniform per wave)
  const int plg = A.glog[s_me];
  const int jl = (tid - A.gbase[s_me]) >> plg;          // proposal slot in a pass
  const int part = (tid - A.gbase[s_me]) & ((1 << plg) - 1);
  const int pch = (((A.Z[s_me] + 3) >> 2) + (1 << plg) - 1) >> plg;  // chunks per part
  const int pz = min(max(A.Z[s_me] - 4 * part * pch, 0), 4 * pch);  // entries of this part
  const int poff = A.off[s_me] + 4 * part * pch;                     // its first column
  const int nq = (task + TP - 1) / TP;                        // passes per round
  int s_beta = 0, s_long = 0;  // shells with the fewest and the most neighbours
  for (int t = 1; t < S; ++t) {
    if (A.Z[t] < A.Z[s_beta]) s_beta = t;
    if (A.Z[t] > A.Z[s_long]) s_long = t;
  }

  Lds<KMAX> T;
  T.carve(smem, task, SP_LDS);
  uint8_t* lat_g = A.lattices + (size_t)l * N;
  uint8_t* sp;
  if constexpr (SP_LDS) {
    sp = T.sp;
    for (int i = tid; i < N; i += nthr) sp[i] = lat_g[i];
  } else {
    sp = lat_g;
  }

  // prev_gamma = R_s(pair counts) (accelerate.cpp:636-640) and constants
  const int32_t* cnt = A.counts + (size_t)l * S * K * K;
  for (int e = tid; e < S * K * K; e += nthr) {
    const int s = e / (K * K), i = (e / K) % K, j = e % K;
    T.G[T.idx(s, i, j)] = hea_r_of_count(P, s, cnt[e]);
    T.tgt[T.idx(s, i, j)] = P.target[s][i][j];
    T.SC[T.idx(s, i, j)].y = -A.delta[s] / (P.conc[i] * P.conc[j]);
  }
  if (tid < K) T.conc[tid] = P.conc[tid];
  if (tid < HEA_MAX_S) {
    T.sh->bmax[tid] = 0ull;
    T.sh->ubmax[tid] = 0ull;
  }
  if (tid == 0) T.sh->gmax = 0u;
  if constexpr (GUIDED) {
    for (int j = tid; j < task; j += nthr) T.gkey[j] = 0ull;
    // absent shells and species keep zero gains
    for (int e = tid; e < HEA_MAX_S * KMAX * 8; e += nthr) T.H[e] = 0.0f;
  }
  if (tid < kMaxWaves) {  // slots of absent waves stay neutral
    T.sh->wcount[tid] = 0;
    T.sh->wmin[tid] = __builtin_inf();
  }
  __syncthreads();
  for (int e = tid; e < S * K * K; e += nthr) {
    const int s = e / (K * K), i = (e / K) % K, j = e % K;
    refresh_entry<KMAX, GUIDED>(A, T, s, i, j);
  }
  __syncthreads();

  if (tid == 0) {
    const double f0 = A.fitness[l];
    T.sh->pf = f0;
    T.sh->pfm = 0;
    T.sh->fest = f0 * f0;
  }
  const uint32_t iter = *A.iter_ptr;
  uint32_t pos = 0;
  int acc = 0;
  long long rounds = 0, accepts = 0, screens = 0;
  bool hit_thr = false;
  // Plain mode: the next round's first draw pass (stream positions pos + tid)
  // is drawn by the waves that wait during the exact stage; wave 1 also draws
  // wave 0's share, handed over through dq.  After an accept its validity
  // ballot is published by the accept's barrier.
  uint32_t nx1 = 0, nx2 = 0;
  bool nxt = false;        // nx1, nx2 hold the pair at pos + tid
  int npk = 0;             // species in HBM: its species a | b << 8 (valid iff they differ)
  int nsq = 0;             // species in HBM: its species before this round's swap
  bool nvalid = false;     // species in LDS: its validity (both when nbal_ok)
  bool nbal_ok = false;    // this wave's ballot of the pass is in wcount already
  // draw lanes per pass: all lanes when species sit in LDS; with species in
  // HBM only as many as a round needs with a wide margin (each draw costs
  // two species loads)
  const int dw = SP_LDS ? nthr : max(64, min(A.dw > 0 ? A.dw : nthr, nthr));
  int2* dq = reinterpret_cast<int2*>(T.cmap);  // [64] (cmap is unused in plain mode)
  // plain mode with records: the probable winner's index and the two rows of
  // its sites, fetched during the exact stage (cmap words 128 and 132..)
  int* pj = reinterpret_cast<int*>(T.cmap + 128);
  int32_t* prow = reinterpret_cast<int32_t*>(T.cmap + 132);  // [2][stride]
  if (!GUIDED && tid == 0) *pj = -1;

  constexpr bool guided = GUIDED;
  // guided mode: hist[k] = neighbour species counts of site k (HeaHistLayout:
  // 8 bytes per site, so the records of many lattices stay cache-resident)
  unsigned long long* hist = guided ? A.hist + (size_t)l * N : nullptr;
  const HeaHistLayout hl = hea_hist_layout(S, A.Z);
  // the screen walk takes its species histograms from the neighbour records
  // instead of gathering every neighbour's species: always in guided mode,
  // which keeps the records anyway (the walk then reads one row per
  // proposal instead of two), and with species in HBM when A.rec is set
  const bool use_rec = GUIDED || (!SP_LDS && !SELF && A.rec != nullptr);
  unsigned long long* rec = use_rec ? (GUIDED ? hist : A.rec + (size_t)l * N) : nullptr;
  // this wave's shell fields in the records (selected without indexing the
  // layout by a variable, which would move it to scratch memory)
  const int r_o = s_me == 0 ? hl.o[0] : s_me == 1 ? hl.o[1] : hl.o[2];
  const int r_w = s_me == 0 ? hl.w[0] : s_me == 1 ? hl.w[1] : hl.w[2];
  if (GUIDED || use_rec) {
    unsigned long long* rb = GUIDED ? hist : rec;
    for (int k = tid; k < N; k += nthr) {
      const int32_t* row = A.nbr + (size_t)k * A.stride;
      unsigned long long m = 0ull;
      if constexpr (GUIDED) {
        for (int s = 0; s < S; ++s)
          for (int q = 0; q < A.Z[s]; ++q) {
            const int c = sp[row[A.off[s] + q]];
            if (c < K - 1) m += 1ull << (c * hl.W + hl.o[s]);
          }
      } else {  // (without a variable index into the layout: see guided_apply_swap)
#pragma unroll
        for (int s = 0; s < HEA_MAX_S; ++s)
          if (s < S)
            for (int q = 0; q < A.Z[s]; ++q) {
              const int c = sp[row[A.off[s] + q]];
              if (c < K - 1) m += 1ull << (c * hl.W + hl.o[s]);
            }
      }
      rb[k] = m;
    }
    __syncthreads();  // every site's record before the first read
  }
  // guided mode: candidate stream key of this lattice and phase
  const uint32_t gk = hea_mix32(hea_mix32(hea_mix32((uint32_t)A.seed ^ 0x6a09e667u) ^
                                          (uint32_t)(A.seed >> 32)) ^
                                (iter * 4u + (A.phase & 3u))) ^
                      hea_mix32((uint32_t)l + 0x3c6ef372u);

  while (acc < A.depth) {
    if constexpr (GUIDED) {
      // ---- guided: slot j takes the Gumbel-max of -gain/(tau*scale) over
      // g_cand hashed candidate pairs (same-species pairs skipped); all lanes
      // score (slot, candidate) items and keep each slot's best in gkey
      const float gs = 2.0f * __uint_as_float(T.sh->gmax);
      const float inv = (float)(1.0 / (A.g_tau * (gs > 0.0f ? (double)gs : 1.0)));
      const int g = A.g_cand, items = task * g;
      if (A.dbg_pred) {  // test capture of the first candidate phase (see HeaMcArgs)
        for (int w = tid; w < items; w += nthr) {
          uint32_t s1, s2, h3;
          guided_pair(gk, pos + (uint32_t)w, pos + (uint32_t)(w - w % g), (uint32_t)N, A.g_blk,
                      &s1, &s2, &h3);
          const int a = sp[s1], b = sp[s2];
          const size_t o = (size_t)l * items + w;
          A.dbg_pred[o] = a != b ? guided_gain<KMAX>(T, hl, K, a, b, hist[s1], hist[s2])
                                 : __builtin_nanf("");
          A.dbg_pairs[2 * o] = (int)s1;
          A.dbg_pairs[2 * o + 1] = (int)s2;
        }
        return;  // no search step: the lattice and its fitness stay as they were
      }
      // four items per lane at a time, so their histogram loads overlap
      for (int w0 = tid; w0 < items; w0 += 4 * nthr) {
        uint32_t s1[4], s2[4], h3[4];
        int sa[4], sb[4];
        bool live[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int w = w0 + q * nthr;
          guided_pair(gk, pos + (uint32_t)w, pos + (uint32_t)(w - w % g), (uint32_t)N, A.g_blk,
                      &s1[q], &s2[q], &h3[q]);
          sa[q] = sp[s1[q]];
          sb[q] = sp[s2[q]];
          live[q] = w < items && sa[q] != sb[q];
        }
        unsigned long long hv[4][2];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          // every access to a lattice's records comes from its own
          // workgroup: workgroup-scope atomics, loads and barriers order them
          hv[q][0] = __hip_atomic_load(hist + s1[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          hv[q][1] = __hip_atomic_load(hist + s2[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const float d = guided_gain<KMAX>(T, hl, K, sa[q], sb[q], hv[q][0], hv[q][1]);
          const float U = ((float)(h3[q] >> 8) + 0.5f) * 0x1p-24f;
          const float sc = -d * inv - __logf(-__logf(U));
          // order-preserving float bits; ties go to the lower candidate
          const uint32_t fb = __float_as_uint(sc);
          const uint32_t ob = (fb & 0x80000000u) ? ~fb : (fb | 0x80000000u);
          const int w = w0 + q * nthr;
          const int j = w / g, c = w - j * g;
          if (live[q])
            atomicMax(&T.gkey[j], ((unsigned long long)ob << 32) | (uint32_t)(0xffffffffu - (uint32_t)c));
          __builtin_amdgcn_sched_barrier(0);  // one item's gain rows at a time
        }
      }
      __syncthreads();
      for (int j = tid; j < task; j += nthr) {
        const unsigned long long key = T.gkey[j];
        T.gkey[j] = 0ull;
        int2 bp = make_int2(-1, -1);
        if (key != 0ull) {
          const int c = (int)(0xffffffffu - (uint32_t)key);
          uint32_t s1, s2, h3;
          guided_pair(gk, pos + (uint32_t)(j * g + c), pos + (uint32_t)(j * g), (uint32_t)N, A.g_blk,
                      &s1, &s2, &h3);
          bp = make_int2((int)s1, (int)s2);
        }
        T.prop[j] = bp;
      }
      pos += (uint32_t)items;
      __syncthreads();
    }
    // ---- draw `task` valid proposals in stream order (accelerate.cpp:649-655)
    int found = guided ? task : 0;
    bool first = true;
    while (found < task) {
      const uint32_t p = pos + tid;  // lanes tid < dw draw
      uint32_t s1 = 0, s2 = 0;
      if (first && nxt) {
        s1 = nx1;
        s2 = nx2;
      } else if (tid < dw) {
        hea_stream_pair(A.seed, iter, A.phase, (uint32_t)l, p, (uint32_t)N, &s1, &s2);
      }
      bool valid;
      unsigned long long bal;
      int ab = npk;  // the pair's species a | b << 8
      if (first && nbal_ok) {
        valid = SP_LDS ? nvalid : (npk & 255) != (npk >> 8);
        bal = __ballot(valid);
      } else {
        if (!SP_LDS && first && nxt)  // loaded during the exact stage, no swap since
          ab = (tid < dw) ? nsq : 0;
        else
          ab = (tid < dw) ? (int)sp[s1] | ((int)sp[s2] << 8) : 0;
        valid = (ab & 255) != (ab >> 8);
        bal = __ballot(valid);
        if (lane == 0) T.sh->wcount[wave] = __popcll(bal);
        __syncthreads();
      }
      first = false;
      const int pre = __popcll(bal & ((1ull << lane) - 1ull));
      int off = 0, tot = 0;
#pragma unroll
      for (int w4 = 0; w4 < kMaxWaves / 4; ++w4) {
        if (4 * w4 >= nwaves) break;  // slots of absent waves hold 0
        const int4 c4 = reinterpret_cast<const int4*>(T.sh->wcount)[w4];
        const int cw[4] = {c4.x, c4.y, c4.z, c4.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          off += (4 * w4 + k < wave) ? cw[k] : 0;
          tot += cw[k];
        }
      }
      const int rank = found + off + pre;
      if (valid && rank < task) {
        T.prop[rank] = make_int2((int)s1, (int)s2);
        if constexpr (!SP_LDS) T.pab[rank] = (uint16_t)ab;
      }
      if (valid && rank == task - 1) T.sh->next_pos = p + 1;
      found += tot;
      pos += dw;
      __syncthreads();
      if (found >= task) pos = T.sh->next_pos;
    }
    nxt = false;
    nbal_ok = false;

    // Guided mode accepts up to g_multi swaps per round: the first as in
    // plain mode, the others from the round's shortlist (guided_more_accepts).
    if constexpr (GUIDED)
      for (int e = tid; e < kMapWords; e += nthr) T.cmap[e] = 0u;
    screens += task;
    // window of the current tables and an empty candidate list, set by the
    // first lane of the shell with the fewest neighbours (its waves finish the
    // screen first); the screen's barrier publishes both
    if (tid == A.gbase[s_beta]) {
      T.sh->beta = screen_beta(A, T);
      T.sh->ncl = 0;
    }
    // ---- screen: walk both neighbour lists of every proposal, keep the walk
    // (the waves of the longest shell finish last: they issue first)
    if (s_me == s_long) __builtin_amdgcn_s_setprio(1);
    for (int q = 0; scr && q < nq; ++q) {
      const int j = jl + q * TP;
      const int2 pr = (j < task) ? T.prop[j] : make_int2(-1, -1);
      double x = 0.0;
      {
        PT MA = 0, MB = 0;
        int c11 = 0, c12 = 0, c22 = 0, c21 = 0, a = 0, b = 0;
        if (pr.x >= 0) {
          if constexpr (!SP_LDS && !GUIDED) {  // the draw loaded them
            const int ab = T.pab[j];
            a = ab & 255;
            b = ab >> 8;
          } else {
            a = sp[pr.x];  // before the walk, so these loads overlap the row loads
            b = sp[pr.y];
          }
          const int32_t* r1 = A.nbr + (size_t)pr.x * A.stride + poff;
          const int32_t* r2 = A.nbr + (size_t)pr.y * A.stride + poff;
          if (use_rec) {
            // histograms from the two records (part 0 of the list only, so
            // the parts' sums below count them once); from the rows only
            // the adjacency: c21 = c12 for a symmetric relation, and no
            // site lists itself
            unsigned long long ra = 0ull, rb = 0ull;
            if (part == 0) {
              ra = __hip_atomic_load(rec + pr.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
              rb = __hip_atomic_load(rec + pr.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
            c12 = (plg == 0) ? count_site<false>(r1, pz, pch, pr.y)
                             : count_site<true>(r1, pz, pch, pr.y);
            c21 = c12;
            if (part == 0) {
              MA = rec_pack<KMAX>(ra, r_o, r_w, hl.W, K, A.Z[s_me]);
              MB = rec_pack<KMAX>(rb, r_o, r_w, hl.W, K, A.Z[s_me]);
            }
          } else if (SELF || plg == 0)  // the whole list (lists of cells with self neighbours are never split)
            walk_pair<KMAX, SELF, false>(r1, r2, pz, pch, sp, pr.x, pr.y, MA, MB, c11, c12, c21,
                                         c22);
          else
            walk_pair<KMAX, SELF, true>(r1, r2, pz, pch, sp, pr.x, pr.y, MA, MB, c11, c12, c21,
                                        c22);
        }
        if (plg > 0) {  // the parts' sums (all lanes of the wave take part)
          MA = part_sum(MA, plg);
          MB = part_sum(MB, plg);
          const uint32_t cc = part_sum((uint32_t)c11 | ((uint32_t)c12 << 8) |
                                           ((uint32_t)c21 << 16) | ((uint32_t)c22 << 24), plg);
          c11 = (int)(cc & 255u);
          c12 = (int)((cc >> 8) & 255u);
          c21 = (int)((cc >> 16) & 255u);
          c22 = (int)(cc >> 24);
        }
        if (part == 0 && pr.x >= 0) {
          const int w = s_me * task + j;
          T.WA[w] = MA;
          T.WB[w] = MB;
          T.WC[w] = walk_pack(c11, c12, c21, c22, a, b);
          x = screen_shell<KMAX, EXACT>(T, EXACT ? KMAX : K, s_me, a, b, MA, MB, c11, c12, c21,
                                        c22);
        }
      }
      if (part == 0 && j < task) T.xs[s_me * task + j] = x;
    }
    if (s_me == s_long) __builtin_amdgcn_s_setprio(0);
    __syncthreads();
    // ---- X = sum_s w_s x_s, each wave's minimum and its candidate list
    const double beta_c = T.sh->beta;  // loaded before the stores below
    double lmin = __builtin_inf();
    double X1 = 0.0;  // this lane's X when the round fits one pass
    bool has1 = false;
    int jmin = -1;    // guided mode: this lane's proposal with the smallest X
    for (int j = tid; j < task; j += nthr) {
      const int2 pr = T.prop[j];
      if (pr.x >= 0) {
        double X = 0.0;
        for (int s = 0; s < S; ++s) X = X + A.weight[s] * T.xs[s * task + j];
        T.xs[j] = X;
        if (X < lmin) jmin = j;
        lmin = fmin(lmin, X);
        X1 = X;
        has1 = true;
      }
    }
    if constexpr (GUIDED) {
      // shortlist for the round's further accepts: each of the nwp waves
      // holding proposals contributes its ceil(64 / nwp) smallest X < 0
      const int nwp = min(nwaves, (task + 63) >> 6);
      const int per = (64 + nwp - 1) / nwp;
      if (A.g_multi > 1 && wave < nwp) {
        double v = (lmin < 0.0) ? lmin : __builtin_inf();
        for (int r = 0; r < per; ++r) {
          const double m = wave_min(v);
          int jr = -1;
          if (m < 0.0) {
            const unsigned long long bal = __ballot(v == m);
            const int src = (int)__builtin_ctzll(bal);
            jr = __builtin_amdgcn_readlane(jmin, src);
            if (lane == src) v = __builtin_inf();
          }
          if (lane == 0 && wave * per + r < 64) T.sl[wave * per + r] = jr;
        }
      }
    }
    lmin = wave_min(lmin);
    if (lane == 0) T.sh->wmin[wave] = lmin;
    {
      // this wave's proposals within 2 beta of the wave's minimum: a superset
      // of the final candidates, since the workgroup's minimum is no larger
      const double wthr = lmin + 2.0 * beta_c;
      if (task <= nthr) {  // one pass: lane tid holds proposal tid's X
        const bool f = has1 && !(X1 > wthr);
        const unsigned long long bal = __ballot(f);
        if (bal) {
          int base = 0;
          if (lane == 0) base = atomicAdd(&T.sh->ncl, __popcll(bal));
          base = __builtin_amdgcn_readfirstlane(base);
          if (f) T.cl[base + __popcll(bal & ((1ull << lane) - 1ull))] = tid;
        }
      } else {
        for (int j0 = 0; j0 < task; j0 += nthr) {
          const int j = j0 + tid;
          bool f = false;
          if (j < task) {
            const int2 pr = T.prop[j];
            f = pr.x >= 0 && !(T.xs[j] > wthr);
          }
          const unsigned long long bal = __ballot(f);
          if (bal) {
            int base = 0;
            if (lane == 0) base = atomicAdd(&T.sh->ncl, __popcll(bal));
            base = __builtin_amdgcn_readfirstlane(base);
            if (f) T.cl[base + __popcll(bal & ((1ull << lane) - 1ull))] = j;
          }
        }
      }
    }
    __syncthreads();

    // ---- exact stage (wave 0): candidates within the window, first strict
    // minimum of F (accelerate.cpp:663-674) and the accept test (:678)
    if (wave == 0) {
      // the exact stage is this workgroup's critical path: issue priority over
      // the waves of other workgroups on the CU until it ends
      __builtin_amdgcn_s_setprio(2);
      const double beta = T.sh->beta;
      const int nsup = T.sh->ncl;
      double xmin = __builtin_inf();
#pragma unroll
      for (int w2 = 0; w2 < kMaxWaves / 2; ++w2) {
        if (2 * w2 >= nwaves) break;  // slots of absent waves hold inf
        const double2 m2 = reinterpret_cast<const double2*>(T.sh->wmin)[w2];
        xmin = fmin(xmin, fmin(m2.x, m2.y));
      }
      exact_stage<KMAX, EXACT>(A, T, sp, lane, task, nsup, xmin, beta, true);
      __builtin_amdgcn_s_setprio(0);
    } else if constexpr (!GUIDED) {
      // with species in HBM their loads for the next pass are issued here
      // too (an accept below replaces the two swapped sites' species)
      hea_stream_pair(A.seed, iter, A.phase, (uint32_t)l, pos + (uint32_t)tid, (uint32_t)N, &nx1,
                      &nx2);
      if constexpr (!SP_LDS)
        if (tid < dw) nsq = (int)sp[nx1] | ((int)sp[nx2] << 8);
      if (EXACT && use_rec && wave == 2) {  // (not in the batch kernels: registers)
        // the probable winner, the candidate with the smallest X (the first
        // in j among equals): its rows for the accept's record updates (the
        // exact stage compacts cl meanwhile; any entry read is a proposal)
        const int ncl = T.sh->ncl;
        double bx = __builtin_inf();
        int bj = 0x7fffffff;
        for (int k = lane; k < ncl; k += 64) {
          const int j = T.cl[k];
          const double x = T.xs[j];
          if (x < bx || (x == bx && j < bj)) {
            bx = x;
            bj = j;
          }
        }
        const double m = wave_min(bx);
        const int js = (int)wave_min((bx == m) ? (double)bj : 2147483647.0);
        if (js < task) {
          const int2 pr = T.prop[js];
          const int rch = A.stride >> 2;
          for (int t = lane; t < 2 * rch; t += 64)
            reinterpret_cast<int4*>(prow)[t] = reinterpret_cast<const int4*>(
                A.nbr + (size_t)(t < rch ? pr.x : pr.y) * A.stride)[t < rch ? t : t - rch];
        }
        if (lane == 0) *pj = js < task ? js : -1;
      }
      if (wave == 1) {  // wave 0's share: species in bits 26.. of the sites (N <= 2^26)
        uint32_t u1, u2;
        hea_stream_pair(A.seed, iter, A.phase, (uint32_t)l, pos + (uint32_t)lane, (uint32_t)N,
                        &u1, &u2);
        if constexpr (!SP_LDS) {
          u1 |= (uint32_t)sp[u1] << 26;
          u2 |= (uint32_t)sp[u2] << 26;
        }
        dq[lane] = make_int2((int)u1, (int)u2);
      }
    }
    __syncthreads();
    const int kind = T.sh->res_kind;
    ++rounds;
    if (!GUIDED && nthr > 64) {
      if (wave == 0) {
        const int2 v = dq[lane];
        nx1 = (uint32_t)v.x & 0x3ffffffu;
        nx2 = (uint32_t)v.y & 0x3ffffffu;
        if constexpr (!SP_LDS) nsq = (int)((uint32_t)v.x >> 26) | ((int)((uint32_t)v.y >> 26) << 8);
      }
      nxt = true;
    }

    // ---- accept iff strictly better (accelerate.cpp:678-689)
    if (kind != 0) {
      const int2 pr = T.prop[T.sh->res_j];
      const int a = T.sh->res_a, b = T.sh->res_b;
      if (kind == 1 && __longlong_as_double((long long)T.sh->res_key) <= A.thr) hit_thr = true;
      acc = 0;
      ++accepts;
      if (tid == 0) {
        sp[pr.x] = (uint8_t)b;
        sp[pr.y] = (uint8_t)a;
      }
      // species at the next pass's pair, loaded before the refresh so the
      // loads overlap it (sp at pr.x, pr.y may be stale here: replaced below)
      int q1 = 0, q2 = 0;
      if constexpr (SP_LDS) {
        if (!GUIDED && nxt) {
          q1 = sp[nx1];
          q2 = sp[nx2];
        }
      } else {  // loaded during the exact stage
        q1 = nsq & 255;
        q2 = nsq >> 8;
      }
      // plain mode with records: the neighbours whose records change, loaded
      // before the refresh so the loads overlap it
      SwapNbrs nbs{{-1, -1, -1}};
      if (!GUIDED && use_rec)
        nbs = (EXACT && *pj == T.sh->res_j) ? swap_nbrs_rows(A, prow, tid, nthr)
                                   : swap_nbrs_load(A, pr, tid, nthr);
      if constexpr (GUIDED)
        guided_apply_swap<KMAX>(A, T, hist, hl, pr, a, b, tid, nthr,
                                swap_nbrs_load(A, pr, tid, nthr));
      for (int e = tid; e < S * K * K; e += nthr) {
        const int s = e / (K * K), i = (e / K) % K, j = e % K;
        refresh_entry<KMAX, GUIDED>(A, T, s, i, j);
      }
      if (!GUIDED && use_rec) guided_apply_swap<KMAX, false>(A, T, rec, hl, pr, a, b, tid, nthr, nbs);
      if (!GUIDED && nxt) {  // species and validity of the next pass after this swap
        const int v1 = (nx1 == (uint32_t)pr.x) ? b : (nx1 == (uint32_t)pr.y) ? a : q1;
        const int v2 = (nx2 == (uint32_t)pr.x) ? b : (nx2 == (uint32_t)pr.y) ? a : q2;
        bool nv;
        if constexpr (SP_LDS) {
          nvalid = v1 != v2;  // every lane draws
          nv = nvalid;
        } else {
          npk = (tid < dw) ? (v1 | (v2 << 8)) : 0;
          nv = (npk & 255) != (npk >> 8);
        }
        const unsigned long long nb = __ballot(nv);
        if (lane == 0) T.sh->wcount[wave] = __popcll(nb);
        nbal_ok = true;
      }
      __syncthreads();
      if constexpr (GUIDED) {
        if (A.g_multi > 1) {
          if (wave == 0) {
            accepts += guided_more_accepts<KMAX, EXACT>(A, T, sp, hist, hl, lane, task,
                                                        A.g_multi - 1, &screens, &hit_thr);
          } else if (wave == 1) {
            // warm this CU's L1 with the rows of the shortlisted proposals'
            // sites, which the further accepts load for their record updates
            const int j = T.sl[lane];
            if (j >= 0) {
              const int2 pr = T.prop[j];
              const int4* r1 = reinterpret_cast<const int4*>(A.nbr + (size_t)pr.x * A.stride);
              const int4* r2 = reinterpret_cast<const int4*>(A.nbr + (size_t)pr.y * A.stride);
              int acc = 0;
              for (int c = 0; c < (A.stride >> 2); c += 2) acc ^= r1[c].x ^ r2[c].x;
              asm volatile("" ::"v"(acc));
            }
          }
          __syncthreads();
        }
      }
      if (hit_thr && tid == 0 && A.first_hit) {
        atomicMin(A.first_hit, (unsigned long long)wall_clock64());
        hit_thr = false;
      }
    } else {
      ++acc;
    }
  }

  double prev_fit = T.sh->pf;
  if (T.sh->pfm == 2 && tid == 0) {  // F of the final gamma (accelerate.cpp:557-569)
    double err = 0.0;
    for (int t = 0; t < S; ++t) {
      double es = 0.0;
      for (int i = 0; i < K * K; ++i) es = es + T.E0[T.idx(t, i / K, i % K)];
      err = err + A.weight[t] * es;
    }
    prev_fit = sqrt(err);
  }
  if constexpr (SP_LDS) {
    for (int i = tid; i < N; i += nthr) lat_g[i] = sp[i];
  }
  if (tid == 0) {
    A.fitness[l] = prev_fit;
    int64_t* st = A.stats + (size_t)l * 4;
    st[0] += rounds;
    st[1] += accepts;
    st[2] += screens;
    st[3] += (long long)pos;
  }
}

template <int KMAX, bool SP_LDS, bool EXACT, bool SELF, bool GUIDED>
// At most 128 VGPRs (4 waves per SIMD): two 6- or 8-wave workgroups share a CU.
__global__ __launch_bounds__(HEA_MC_MAX_THREADS, 4) void k_local_mc(HeaMcArgs A) {
  extern __shared__ __align__(16) unsigned char smem[];
  mc_lattice<KMAX, SP_LDS, EXACT, SELF, GUIDED>(A, blockIdx.x, smem);
}

// Persistent local search over many problems: every workgroup takes
// (problem, lattice) items from a device queue, longest first, until the
// queue is empty.  B.probs[p] holds problem p's complete launch arguments
// (its seed is the batch seed + p).
template <int KMAX, bool SP_LDS, bool GUIDED>
__global__ __launch_bounds__(HEA_MC_MAX_THREADS, 4) void k_local_mc_batch(HeaMcBatchArgs B) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int item_sh;
  // the part of this workgroup's XCD first (its problems' neighbour tables
  // stay in that XCD's L2), then the others' when it runs dry; every part is
  // tried at most once after it is found empty, so the loop ends
  const int home = blockIdx.x % HEA_XCDS;
  int tried = 0;  // parts found empty (thread 0)
  for (;;) {
    if (threadIdx.x == 0) {
      int it = -1;
      while (tried < HEA_XCDS) {
        const int x = (home + tried) % HEA_XCDS;
        const uint32_t k = atomicAdd(&B.queue[x], 1u);
        if (k < (uint32_t)(B.xoff[x + 1] - B.xoff[x])) {
          it = B.xoff[x] + (int)k;
          break;
        }
        ++tried;
      }
      item_sh = it;
    }
    __syncthreads();
    const int item = __builtin_amdgcn_readfirstlane(item_sh);
    __syncthreads();
    if (item < 0) break;
    const int2 it = B.items[item];
    const HeaMcArgs& A = B.probs[it.x];
    mc_lattice<KMAX, SP_LDS, false, false, GUIDED>(A, it.y, smem);
    __syncthreads();
  }
}

template <int KMAX, bool SP_LDS, bool EXACT, bool SELF, bool GUIDED>
const void* mc_fn() {
  return reinterpret_cast<const void*>(k_local_mc<KMAX, SP_LDS, EXACT, SELF, GUIDED>);
}

template <int KMAX, bool EXACT, bool SELF, bool GUIDED>
const void* mc_fn(bool sp_lds) {
  return sp_lds ? mc_fn<KMAX, true, EXACT, SELF, GUIDED>() : mc_fn<KMAX, false, EXACT, SELF, GUIDED>();
}

// Kernel variant: exact-K instantiations for the common case, a guarded
// K <= KMAX one for K = 7 and mixed batches, and a self-neighbour one for
// cells so small that a site lists itself.  Guided proposals exist for
// K <= 6 without self neighbours (the host checks this).
const void* mc_fn_for(int K, bool sp_lds, bool exact_ok, bool self, bool guided) {
  if (guided) {
    if (self || K > 6) return nullptr;
    if (exact_ok) {
      switch (K) {
        case 2: return mc_fn<2, true, false, true>(sp_lds);
        case 3: return mc_fn<3, true, false, true>(sp_lds);
        case 4: return mc_fn<4, true, false, true>(sp_lds);
        case 5: return mc_fn<5, true, false, true>(sp_lds);
        case 6: return mc_fn<6, true, false, true>(sp_lds);
        default: break;
      }
    }
    return mc_fn<6, false, false, true>(sp_lds);
  }
  if (self) return mc_fn<8, false, true, false>(sp_lds);
  if (exact_ok) {
    switch (K) {
      case 2: return mc_fn<2, true, false, false>(sp_lds);
      case 3: return mc_fn<3, true, false, false>(sp_lds);
      case 4: return mc_fn<4, true, false, false>(sp_lds);
      case 5: return mc_fn<5, true, false, false>(sp_lds);
      case 6: return mc_fn<6, true, false, false>(sp_lds);
      case 8: return mc_fn<8, true, false, false>(sp_lds);
      default: break;
    }
  }
  if (K <= 6) return mc_fn<6, false, false, false>(sp_lds);
  if (K <= 8) return mc_fn<8, false, false, false>(sp_lds);
  return nullptr;
}

template <int KMAX, bool GUIDED>
const void* mc_batch_fn(bool sp_lds) {
  return sp_lds ? reinterpret_cast<const void*>(k_local_mc_batch<KMAX, true, GUIDED>)
                : reinterpret_cast<const void*>(k_local_mc_batch<KMAX, false, GUIDED>);
}

// Batch kernels exist for table widths 5, 6 and 8 (guided: 5 and 6, its
// packed neighbour histograms hold 6 species); width 5 keeps the tables of
// 108000-site, 5-species problems small enough for their species to fit in
// LDS.
const void* mc_batch_fn_for(int kmax, bool sp_lds, bool guided) {
  if (guided) {
    if (kmax == 5) return mc_batch_fn<5, true>(sp_lds);
    return kmax == 6 ? mc_batch_fn<6, true>(sp_lds) : nullptr;
  }
  if (kmax == 5) return mc_batch_fn<5, false>(sp_lds);
  if (kmax == 6) return mc_batch_fn<6, false>(sp_lds);
  return kmax == 8 ? mc_batch_fn<8, false>(sp_lds) : nullptr;
}

}  // namespace

// Raises the batch kernels' dynamic-LDS limit (plain and, where it exists,
// guided); call before launching or capturing.
hipError_t hea_prepare_local_mc_batch(int kmax, bool sp_lds, size_t lds) {
  for (int g = 0; g < 2; ++g) {
    const void* fn = mc_batch_fn_for(kmax, sp_lds, g != 0);
    if (!fn) continue;
    const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t hea_launch_local_mc_batch(const HeaMcBatchArgs& b, int kmax, bool sp_lds, bool guided,
                                     int grid, int threads, size_t lds, hipStream_t st) {
  const void* fn = mc_batch_fn_for(kmax, sp_lds, guided);
  if (!fn) return hipErrorInvalidValue;
  void* args[] = {const_cast<HeaMcBatchArgs*>(&b)};
  return hipLaunchKernel(fn, dim3(grid), dim3(threads), args, lds, st);
}

// Screen lanes of a launch with a.tp proposal slots per pass: shell s gets
// 2^glog[s] lanes per slot, each walking one part of its neighbour list.  A
// shell is split only when its list has at least twice the chunks (4 entries
// each) of the next-longest list, into parts no shorter than that list: the
// long shell then stops keeping the others waiting, and shorter gains do not
// pay for the extra lanes (measured: FCC 12/6/24 gains, BCC 8/6/12 loses).
// Tiny cells whose sites list themselves (self) keep whole lists, and so do
// walks that read neighbour records (a.rec): their lanes only scan a1's
// row, and unsplit lists measured 3 % faster at FCC 50^3.  Returns the lanes.
int hea_mc_split(HeaMcArgs& a, bool self) {
  int nch[HEA_MAX_S], lg[HEA_MAX_S] = {0, 0, 0};
  for (int s = 0; s < a.S; ++s) nch[s] = std::max(1, (a.Z[s] + 3) >> 2);
  for (int s = 0; s < a.S && !self && !a.rec; ++s) {
    int next = 0;  // longest other list
    for (int t = 0; t < a.S; ++t)
      if (t != s) next = std::max(next, nch[t]);
    while (next > 0 && lg[s] < 2 && nch[s] >= (2 * next) << lg[s]) ++lg[s];
  }
  for (;;) {  // fit the workgroup: halve the widest split until it does
    int lanes = 0, wide = 0;
    for (int s = 0; s < a.S; ++s) {
      lanes += a.tp << lg[s];
      if (lg[s] > lg[wide]) wide = s;
    }
    if (lanes <= HEA_MC_MAX_THREADS || lg[wide] == 0) break;
    --lg[wide];
  }
  int g = 0;
  for (int s = 0; s < HEA_MAX_S; ++s) {
    a.glog[s] = s < a.S ? lg[s] : 0;
    a.gbase[s] = g;
    if (s < a.S) g += a.tp << lg[s];
  }
  for (int s = a.S; s <= HEA_MAX_S; ++s) a.gbase[s] = g;
  return g;
}

// Draw lanes per pass when species are read from HBM (every draw then costs
// two species loads): enough that one pass finds `task` valid draws with a
// wide margin, p = 1 - sum_i c_i^2 being the chance that two random sites
// hold different species (about 13 standard deviations to spare for 256
// proposals of six species); a short pass only costs another pass.
int hea_mc_draw_lanes(int task, const double* conc, int K, int lanes) {
  double q = 0.0;
  for (int i = 0; i < K; ++i) q += conc[i] * conc[i];
  const double p = std::max(1e-3, 1.0 - q);
  const double need = (task + 4.0 * std::sqrt(task / p) + 16.0) / p;
  const int dw = (int)std::min(1e9, std::ceil(need / 64.0)) * 64;
  return std::max(64, std::min(dw, lanes));
}

// Table width of the kernel chosen for K species (see mc_fn_for).
int hea_mc_kmax(int K, bool self) {
  if (self || K == 7) return 8;
  return K <= 6 ? K : 8;
}

// Proposals evaluated per pass (lanes per shell): a multiple of 64, at most
// HEA_MC_MAX_TP, so S * TP <= HEA_MC_MAX_THREADS lanes.
int hea_mc_tp(int task, int S) {
  const int want = ((task + 63) / 64) * 64;
  const int cap = std::min(HEA_MC_MAX_TP, (HEA_MC_MAX_THREADS / S) / 64 * 64);
  return std::min(want, cap);
}

// LDS bytes of one workgroup (the pass width does not change the layout).
size_t hea_mc_lds_bytes(int kmax, int task, int tp, int n_sites, bool sp_lds) {
  (void)tp;
  switch (kmax) {
    case 2: return Lds<2>::bytes(task, n_sites, sp_lds);
    case 3: return Lds<3>::bytes(task, n_sites, sp_lds);
    case 4: return Lds<4>::bytes(task, n_sites, sp_lds);
    case 5: return Lds<5>::bytes(task, n_sites, sp_lds);
    case 6: return Lds<6>::bytes(task, n_sites, sp_lds);
    default: return Lds<8>::bytes(task, n_sites, sp_lds);
  }
}

// Raises the kernel's dynamic-LDS limit (plain and guided variants); call
// before launching (and before any graph capture that contains the launch).
hipError_t hea_prepare_local_mc(int K, bool sp_lds, bool self, size_t lds) {
  const void* fn = mc_fn_for(K, sp_lds, true, self, false);
  if (!fn) return hipErrorInvalidValue;
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  const void* gfn = mc_fn_for(K, sp_lds, true, self, true);
  if (e == hipSuccess && gfn)
    e = hipFuncSetAttribute(gfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  return e;
}

hipError_t hea_launch_local_mc(const HeaMcArgs& a, int K, bool sp_lds, bool self, size_t lds,
                               hipStream_t st) {
  const void* fn = mc_fn_for(K, sp_lds, true, self, a.hist != nullptr);
  if (!fn) return hipErrorInvalidValue;
  void* args[] = {const_cast<HeaMcArgs*>(&a)};
  return hipLaunchKernel(fn, dim3(a.L), dim3(a.gbase[a.S]), args, lds, st);
}
