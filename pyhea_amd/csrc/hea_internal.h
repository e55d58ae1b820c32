// Internal layout shared by the HIP kernels and the host side of the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "hea_stream.h"

#define HEA_MAX_K 8        // species per lattice handled by the kernels
#define HEA_MAX_S 3        // neighbour shells (the reference builds 3)
#define HEA_MAX_Z 31       // members per shell (5-bit packed counters)
#define HEA_MC_MAX_TP 256   // proposals per pass (lanes per shell) at most
#define HEA_MC_MAX_THREADS 1024  // local-search workgroup size at most
#define HEA_MAX_TASK 2048  // proposals per round (parallel_task)
#define HEA_MAX_SEG 160    // binade segments of one shell's R_s(k) table

// Per-problem constants, resident in HBM and read through the scalar cache.
//
// gamma entries of a fresh recount equal R_s(k): the k-fold sequential FP64
// sum of delta_s = 1/(Z_s*N) (pyhea/cpp/accelerate.cpp:160, 395).  R_s is
// piecewise linear: within a run of equal increments R_s(k) = v0 + (k-k0)*inc
// exactly, so it is stored as segments (k0, v0, inc) found by walking the
// sum once on the host.
struct HeaConst {
  int32_t N, K, S, stride;       // stride: ints per neighbour row (multiple of 4)
  int32_t Z[HEA_MAX_S];          // shell sizes
  int32_t off[HEA_MAX_S];        // first column of each shell in a row (multiple of 4)
  int32_t nseg[HEA_MAX_S];
  int32_t self_nbr;              // 1 if some site lists itself (tiny cells)
  double delta[HEA_MAX_S];       // fl(1.0/(Z_s*N))
  double weight[HEA_MAX_S];
  double conc[HEA_MAX_K];           // fl(count_i/N)
  int32_t prefix[HEA_MAX_K + 1];     // prefix sums of the species counts
  double dd[HEA_MAX_K][HEA_MAX_K];  // fl(fl(c_i/N)*fl(c_j/N)), accelerate.cpp:562
  double target[HEA_MAX_S][HEA_MAX_K][HEA_MAX_K];
  int64_t seg_k0[HEA_MAX_S][HEA_MAX_SEG];
  double seg_v0[HEA_MAX_S][HEA_MAX_SEG];
  double seg_inc[HEA_MAX_S][HEA_MAX_SEG];
};

// Arguments of one local-search launch over a population of L lattices.
struct HeaMcArgs {
  const HeaConst* pc;
  const int32_t* nbr;       // [N][stride], per-shell lists ascending, padding -1
  uint8_t* lattices;        // [L][N] in/out
  double* fitness;          // [L] in/out (stored fitness = prev_fit)
  const int32_t* counts;    // [L][S][K][K] integer pair counts of the input lattices
  int64_t* stats;           // [L][4] += rounds, accepts, evals, draws
  const uint32_t* iter_ptr; // stream iteration (device memory, so graphs can replay)
  unsigned long long* first_hit;  // wall clock of the first fitness <= thr (atomicMin)
  uint64_t seed;
  int32_t L, task, depth, tp;   // tp: proposal slots per pass
  int32_t dw;                   // draw lanes per pass with species in HBM (0: every lane)
  // screen lanes (hea_mc_split): shell s's group starts at lane gbase[s] and
  // walks its list in 2^glog[s] parts; gbase[S] = lanes of the workgroup
  int32_t gbase[HEA_MAX_S + 1];
  int32_t glog[HEA_MAX_S];
  uint32_t phase;
  double thr;
  // guided proposals (hist == nullptr: plain keyed stream)
  unsigned long long* hist;     // [L][N] per-site neighbour species counts (HeaHistLayout)
  int32_t g_cand;               // candidate pairs per proposal
  double g_tau;                 // softmax temperature in units of the gain scale
  int32_t g_multi;              // accepts per round at most
  int32_t g_blk;                // candidate sites: 0 uniform, b > 0 from two blocks of
                                // 2^b consecutive sites per proposal slot
  // test capture of the first candidate phase (nullptr: none): when set, the
  // search records every (slot, candidate) pair [L][task*g_cand][2] and its
  // predicted change of F^2 [L][task*g_cand] (NaN for same-species pairs),
  // then returns without searching
  int32_t* dbg_pairs;
  float* dbg_pred;
  // neighbour records for the plain screen walk when species are read from
  // HBM (nullptr: the walk gathers every neighbour's species).  Guided mode
  // always walks with its own records (hist).
  unsigned long long* rec;      // [L][N] (HeaHistLayout)
  // problem scalars, copied from HeaConst so they live in the kernarg segment
  int32_t N, K, S, stride;
  int32_t Z[HEA_MAX_S], off[HEA_MAX_S];
  double delta[HEA_MAX_S], weight[HEA_MAX_S];
};

// One launch of the persistent local search over many problems: probs[p]
// holds problem p's complete launch arguments; items[i] = (problem,
// lattice).  The items are split into HEA_XCDS parts, part x =
// items[xoff[x], xoff[x+1]), one per XCD (each problem's lattices in one
// part, longest first within a part); a workgroup takes items from its own
// XCD's part, then from the others.
#define HEA_XCDS 8  // XCDs of an MI355X; workgroups are dispatched round-robin over them
struct HeaMcBatchArgs {
  const HeaMcArgs* probs;
  const int2* items;
  int32_t n_items;
  int32_t xoff[HEA_XCDS + 1];
  uint32_t* queue;  // [HEA_XCDS] next item of each part, zeroed before each launch
};

// Per-site neighbour record: the species counts of a site's neighbours, all
// shells in one 64-bit word.  The count of species x < K-1 in shell s is
// the field at bit x*W + o[s] of width w[s] (the bit width of Z_s; W = sum of
// the widths); species K-1's count is Z_s minus the others.  Usable when
// (K-1)*W <= 64 (FCC 12/6/24 needs 12 bits per species, BCC 8/6/12 11).
// Guided mode scores candidates from them and its screen walk takes its
// species histograms from them; so does the plain walk when species are
// read from HBM (exact:
// the neighbour relation is symmetric, checked when a problem is created,
// so a swap changes exactly the records of the two sites' neighbours).
struct HeaHistLayout {
  int w[HEA_MAX_S], o[HEA_MAX_S], W;
};
__host__ __device__ inline HeaHistLayout hea_hist_layout(int S, const int32_t* Z) {
  HeaHistLayout h;
  h.W = 0;
  for (int s = 0; s < HEA_MAX_S; ++s) {
    int b = 0;
    while (s < S && (1 << b) <= Z[s]) ++b;
    h.w[s] = b;
    h.o[s] = h.W;
    h.W += b;
  }
  return h;
}

// True when a problem's records fit one 64-bit word and the walk may use
// them (cells whose sites list themselves keep the gathering walk).
inline bool hea_records_fit(int K, int S, const int32_t* Z, bool self) {
  return !self && K <= 6 && (K - 1) * hea_hist_layout(S, Z).W <= 64;
}

// R_s(k) from the binade segments (exact: the host verified every k).
__device__ __forceinline__ double hea_r_of_count(const HeaConst& P, int s, int64_t k) {
  int q = 0;
  while (q + 1 < P.nseg[s] && P.seg_k0[s][q + 1] <= k) ++q;
  return fma((double)(k - P.seg_k0[s][q]), P.seg_inc[s][q], P.seg_v0[s][q]);
}

// Launchers implemented in hea_mc.hip.
size_t hea_mc_lds_bytes(int kmax, int task, int tp, int n_sites, bool sp_lds);
int hea_mc_kmax(int K, bool self);
int hea_mc_tp(int task, int S);
int hea_mc_split(HeaMcArgs& a, bool self);
int hea_mc_draw_lanes(int task, const double* conc, int K, int lanes);
hipError_t hea_prepare_local_mc(int K, bool sp_lds, bool self, size_t lds);
hipError_t hea_launch_local_mc(const HeaMcArgs& a, int K, bool sp_lds, bool self, size_t lds,
                               hipStream_t st);
hipError_t hea_prepare_local_mc_batch(int kmax, bool sp_lds, size_t lds);
hipError_t hea_launch_local_mc_batch(const HeaMcBatchArgs& b, int kmax, bool sp_lds, bool guided,
                                     int grid, int threads, size_t lds, hipStream_t st);
