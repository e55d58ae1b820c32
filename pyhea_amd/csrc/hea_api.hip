// Host side of the C ABI (include/pyhea_hcs.h) and the kernels around the
// local search: keyed permutations, pair-count recount, recount fitness,
// stable population merge, and the per-iteration HIP graph.
//
// Reference counterparts (pyhea/cpp/accelerate.cpp):
//   generate_random_lattices     218-252  -> k_permute + k_recount + k_fitness
//   calculate_fitness_of_solutions 118-180 -> k_recount + k_fitness
//   sort_lattices_by_fitness       37-64  -> k_rank + k_gather
//   calculate_best_lattices       305-330 -> k_rank (old ++ new) + k_gather
//   run_local_parallel_hcs        713-763 -> hea_session_* (one graph/iteration)
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <utility>
#include <functional>
#include <string>
#include <vector>

#include "../../include/pyhea_hcs.h"
#include "hea_internal.h"
#include "hea_sort.h"

// Workspace of hea_local_mc_device_noalloc: the local search's iteration
// word, per-lattice counters and pair counts live here instead of in
// buffers the library allocates.  Each device holds one copy for the life of
// the process; every call on a device shares it.  Host code never passes
// `scratch` itself to a kernel (on the host it names a stand-in, not device
// memory): it uses the address hipGetSymbolAddress returned for the device.
static __device__ __align__(256) unsigned char scratch[HEA_NOALLOC_SCRATCH_BYTES];

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HEA_HIP(call)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(HEA_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));       \
  } while (0)

// Device address of `scratch`, one per device, resolved at problem creation
// with that device current.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

int resolve_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return fail(HEA_EUNSUPPORTED, "device index >= 64");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return HEA_OK;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return fail(HEA_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_ptr[device] = static_cast<unsigned char*>(ptr);
  return HEA_OK;
}

unsigned char* scratch_of(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_ptr[device];
}

// ------------------------------------------------------------------ kernels --

// lattice[l][i] = sorted_multiset[perm_l(i)] for i = i0, i0 + di, ... < N.
__device__ __forceinline__ void permute_sites(const HeaConst& P, uint8_t* __restrict__ out,
                                              uint64_t seed, uint32_t iter, uint32_t phase,
                                              int slot, int l, int i0, int di) {
  if (i0 >= P.N) return;
  const hea_perm_key key = hea_perm_make_key(seed, iter, phase, (uint32_t)slot, (uint32_t)P.N);
  for (int i = i0; i < P.N; i += di) {
    const uint32_t q = hea_perm_apply(&key, (uint32_t)i);
    out[(size_t)l * P.N + i] = (uint8_t)hea_species_of_rank(P.prefix, P.K, q);
  }
}

__global__ void k_permute(const HeaConst* __restrict__ pc, uint8_t* __restrict__ out,
                          uint64_t seed, const uint32_t* __restrict__ iter_ptr, uint32_t phase,
                          int slot0) {
  const int l = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  permute_sites(*pc, out, seed, *iter_ptr, phase, slot0 + l, l, i, gridDim.x * blockDim.x);
}

__global__ void k_zero_i32(int32_t* __restrict__ p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

constexpr int kRecountThreads = 256;
constexpr int kRecountSites = 2048;  // sites per workgroup (lattices beyond LDS)
constexpr int kRecountLdsThreads = 1024;

// Species histogram (8-bit fields, Z <= 31) of one shell's neighbour list of
// NC = ceil(z/4) 16-byte chunks, all loaded before any is used.  Lists are
// padded to a multiple of 4 entries.
template <int NC>
__device__ __forceinline__ uint64_t shell_hist(const int32_t* __restrict__ row, int z,
                                               const uint8_t* __restrict__ sp) {
  int4 u[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) u[c] = reinterpret_cast<const int4*>(row)[c];
  uint64_t m = 0;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const int nb[4] = {u[c].x, u[c].y, u[c].z, u[c].w};
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const bool ok = (c < NC - 1) || (4 * c + t < z);  // uniform
      const uint64_t f = 1ull << (8 * sp[ok ? nb[t] : 0]);
      m += ok ? f : 0ull;
    }
  }
  return m;
}

__device__ __forceinline__ uint64_t shell_hist(const int32_t* __restrict__ row, int z,
                                               const uint8_t* __restrict__ sp) {
  switch ((z + 3) >> 2) {  // uniform
    case 1: return shell_hist<1>(row, z, sp);
    case 2: return shell_hist<2>(row, z, sp);
    case 3: return shell_hist<3>(row, z, sp);
    case 4: return shell_hist<4>(row, z, sp);
    case 5: return shell_hist<5>(row, z, sp);
    case 6: return shell_hist<6>(row, z, sp);
    case 7: return shell_hist<7>(row, z, sp);
    default: return shell_hist<8>(row, z, sp);
  }
}

// Integer pair counts n[s][a][b] of every lattice, added to `counts` (which
// the caller zeroes): the count form of calculate_sro_coefficient
// (accelerate.cpp:370-399).  Several workgroups per lattice; species read
// from HBM.  Used for lattices that do not fit in LDS.
__global__ __launch_bounds__(kRecountThreads) void k_recount(
    const HeaConst* __restrict__ pc, const int32_t* __restrict__ nbr,
    const uint8_t* __restrict__ lattices, int32_t* __restrict__ counts) {
  __shared__ int32_t hist[HEA_MAX_S * HEA_MAX_K * HEA_MAX_K];
  const HeaConst& P = *pc;
  const int l = blockIdx.y;
  const int K = P.K, S = P.S;
  for (int e = threadIdx.x; e < S * K * K; e += blockDim.x) hist[e] = 0;
  __syncthreads();
  const uint8_t* lat = lattices + (size_t)l * P.N;
  const int i0 = blockIdx.x * kRecountSites;
  const int i1 = min(P.N, i0 + kRecountSites);
  for (int i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    const int a = lat[i];
    const int32_t* row = nbr + (size_t)i * P.stride;
    for (int s = 0; s < S; ++s) {
      const uint64_t m = shell_hist(row + P.off[s], P.Z[s], lat);
      for (int x = 0; x < K; ++x) {
        const int f = (int)((m >> (8 * x)) & 255u);
        if (f) atomicAdd(&hist[(s * K + a) * K + x], f);
      }
    }
  }
  __syncthreads();
  int32_t* dst = counts + (size_t)l * S * K * K;
  for (int e = threadIdx.x; e < S * K * K; e += blockDim.x)
    if (hist[e]) atomicAdd(&dst[e], hist[e]);
}

// The same with one workgroup per lattice and the lattice's species staged
// in LDS; per-wave histogram tables, then one store per entry (no zeroing of
// `counts` needed).
__device__ __forceinline__ void recount_lds_one(const HeaConst& P, const int32_t* __restrict__ nbr,
                                                const uint8_t* __restrict__ lattices,
                                                int32_t* __restrict__ counts, int l,
                                                unsigned char* smem) {
  constexpr int TAB = HEA_MAX_S * HEA_MAX_K * HEA_MAX_K;
  const int K = P.K, S = P.S, N = P.N, SKK = S * K * K;
  const int tid = threadIdx.x, wave = tid >> 6, nw = blockDim.x >> 6;
  int32_t* wh = reinterpret_cast<int32_t*>(smem);  // [nw][TAB]
  uint8_t* sp = smem + sizeof(int32_t) * TAB * (kRecountLdsThreads / 64);
  const uint8_t* lat = lattices + (size_t)l * N;
  for (int i = tid; i < N; i += blockDim.x) sp[i] = lat[i];
  for (int e = tid; e < nw * TAB; e += blockDim.x) wh[e] = 0;
  __syncthreads();
  int32_t* my = wh + wave * TAB;
  for (int i = tid; i < N; i += blockDim.x) {
    const int a = sp[i];
    const int32_t* row = nbr + (size_t)i * P.stride;
    for (int s = 0; s < S; ++s) {
      const uint64_t m = shell_hist(row + P.off[s], P.Z[s], sp);
      for (int x = 0; x < K; ++x) {
        const int f = (int)((m >> (8 * x)) & 255u);
        if (f) atomicAdd(&my[(s * K + a) * K + x], f);
      }
    }
  }
  __syncthreads();
  int32_t* dst = counts + (size_t)l * SKK;
  for (int e = tid; e < SKK; e += blockDim.x) {
    int sum = 0;
    for (int w = 0; w < nw; ++w) sum += wh[w * TAB + e];
    dst[e] = sum;
  }
}

__global__ __launch_bounds__(kRecountLdsThreads) void k_recount_lds(
    const HeaConst* __restrict__ pc, const int32_t* __restrict__ nbr,
    const uint8_t* __restrict__ lattices, int32_t* __restrict__ counts) {
  extern __shared__ __align__(16) unsigned char smem[];
  recount_lds_one(*pc, nbr, lattices, counts, blockIdx.x, smem);
}

size_t recount_lds_bytes(int n) {
  return sizeof(int32_t) * HEA_MAX_S * HEA_MAX_K * HEA_MAX_K * (kRecountLdsThreads / 64) + n;
}

// Fitness of freshly recounted lattices: gamma = R_s(n), then the
// reference's sum (accelerate.cpp:165-178; std::pow(x, 2) there is x*x).
__device__ __forceinline__ void fitness_one(const HeaConst& P, const int32_t* __restrict__ counts,
                                            double* __restrict__ fit, int L, int l) {
  if (l >= L) return;
  const int K = P.K, S = P.S;
  const int32_t* c = counts + (size_t)l * S * K * K;
  double error = 0.0;
  for (int s = 0; s < S; ++s) {
    double es = 0.0;
    for (int i = 0; i < K; ++i)
      for (int j = 0; j < K; ++j) {
        const double g = hea_r_of_count(P, s, c[(s * K + i) * K + j]);
        const double sro = 1.0 - g / (P.conc[i] * P.conc[j]);
        const double df = sro - P.target[s][i][j];
        es = es + df * df;
      }
    error = error + P.weight[s] * es;
  }
  fit[l] = sqrt(error);
}

__global__ void k_fitness(const HeaConst* __restrict__ pc, const int32_t* __restrict__ counts,
                          double* __restrict__ fit, int L) {
  fitness_one(*pc, counts, fit, L, blockIdx.x * blockDim.x + threadIdx.x);
}

// One population of every problem of a batch (problem q in the last grid
// dimension): its lattices, their pair counts and fitness, and the seed of
// its keyed permutations.
struct HeaPopDesc {
  const HeaConst* pc;
  const int32_t* nbr;
  uint8_t* lat;
  int32_t* counts;
  double* fit;
  uint64_t seed;
};

// k_permute over every problem; blockIdx.x of gridDim.x takes every
// gridDim.x-th chunk of 256 sites.
__global__ void k_permute_batch(const HeaPopDesc* __restrict__ pd,
                                const uint32_t* __restrict__ iter_ptr, uint32_t phase) {
  const HeaPopDesc d = pd[blockIdx.z];
  const int l = blockIdx.y;
  permute_sites(*d.pc, d.lat, d.seed, *iter_ptr, phase, l, l,
                blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

__global__ __launch_bounds__(kRecountLdsThreads) void k_recount_lds_batch(
    const HeaPopDesc* __restrict__ pd) {
  extern __shared__ __align__(16) unsigned char smem[];
  const HeaPopDesc d = pd[blockIdx.y];
  recount_lds_one(*d.pc, d.nbr, d.lat, d.counts, blockIdx.x, smem);
}

__global__ void k_fitness_batch(const HeaPopDesc* __restrict__ pd, int L) {
  const HeaPopDesc d = pd[blockIdx.y];
  fitness_one(*d.pc, d.counts, d.fit, L, blockIdx.x * blockDim.x + threadIdx.x);
}

// Rank by (fitness, position) over fa[0..na) ++ fb[0..nb); writes
// perm[rank] = position for ranks below `keep`.  This is the reference's
// order (std::sort of the positions by fitness, accelerate.cpp:37-64) as
// long as no two kept keys are equal; a kept key that equals another raises
// *tie, and k_tie_sort then redoes the order with the reference's algorithm.
__device__ __forceinline__ void rank_one(const double* __restrict__ fa, int na,
                                         const double* __restrict__ fb, int nb, int keep,
                                         int32_t* __restrict__ perm, uint32_t* __restrict__ tie,
                                         int i) {
  const int n = na + nb;
  if (i >= n) return;
  const double fi = i < na ? fa[i] : fb[i - na];
  int r = 0;
  bool eq = false;
  for (int j = 0; j < n; ++j) {
    const double fj = j < na ? fa[j] : fb[j - na];
    r += (fj < fi) || (fj == fi && j < i);
    eq |= (fj == fi) && (j != i);
  }
  if (r < keep) {
    perm[r] = i;
    if (eq) *tie = 1u;
  }
}

__global__ void k_rank(const double* __restrict__ fa, int na, const double* __restrict__ fb,
                       int nb, int keep, int32_t* __restrict__ perm, uint32_t* __restrict__ tie) {
  rank_one(fa, na, fb, nb, keep, perm, tie, blockIdx.x * blockDim.x + threadIdx.x);
}

// Population order on equal keys: when k_rank saw a tie, one lane sorts the
// positions 0..n-1 with the reference library's std::sort algorithm
// (hea_sort.h), in LDS when `lds` holds the keys and positions (12 bytes
// each), else in place in perm (n entries) reading the keys from HBM; then
// clears *tie.  Without a tie the kernel returns at once.
__device__ __forceinline__ void tie_sort_one(const double* __restrict__ fa, int na,
                                             const double* __restrict__ fb, int nb, int keep,
                                             int32_t* __restrict__ perm,
                                             uint32_t* __restrict__ tie, int lds,
                                             unsigned char* smem) {
  if (__builtin_amdgcn_readfirstlane((int)*tie) == 0) return;
  const int n = na + nb;
  const bool in_lds = lds > 0;
  double* key = reinterpret_cast<double*>(smem);
  int* idx = in_lds ? reinterpret_cast<int*>(smem + sizeof(double) * (size_t)n) : perm;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    if (in_lds) key[i] = i < na ? fa[i] : fb[i - na];
    idx[i] = i;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const hea_sort_keys K = in_lds ? hea_sort_keys{key, nullptr, (int64_t)n}
                                   : hea_sort_keys{fa, fb, (int64_t)na};
    hea_sort_like_std(idx, n, K);
  }
  __syncthreads();
  if (in_lds)
    for (int r = threadIdx.x; r < keep; r += blockDim.x) perm[r] = idx[r];
  if (threadIdx.x == 0) *tie = 0u;
}

__global__ __launch_bounds__(256) void k_tie_sort(const double* __restrict__ fa, int na,
                                                  const double* __restrict__ fb, int nb, int keep,
                                                  int32_t* __restrict__ perm,
                                                  uint32_t* __restrict__ tie, int lds) {
  extern __shared__ __align__(16) unsigned char smem[];
  tie_sort_one(fa, na, fb, nb, keep, perm, tie, lds, smem);
}

// dst row r = (perm[r] < na ? srcA[perm[r]] : srcB[perm[r]-na]); chunk
// blockIdx.x of gridDim.x.
__device__ __forceinline__ void gather_row(const uint8_t* __restrict__ srcA,
                                           const double* __restrict__ fA, int na,
                                           const uint8_t* __restrict__ srcB,
                                           const double* __restrict__ fB,
                                           const int32_t* __restrict__ perm,
                                           uint8_t* __restrict__ dst, double* __restrict__ fdst,
                                           int N, int r) {
  const int p = perm[r];
  const uint8_t* src = p < na ? srcA + (size_t)p * N : srcB + (size_t)(p - na) * N;
  uint8_t* d = dst + (size_t)r * N;
  if ((N & 15) == 0) {
    const uint4* s4 = reinterpret_cast<const uint4*>(src);
    uint4* d4 = reinterpret_cast<uint4*>(d);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N / 16; i += gridDim.x * blockDim.x)
      d4[i] = s4[i];
  } else {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
      d[i] = src[i];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) fdst[r] = p < na ? fA[p] : fB[p - na];
}

__global__ void k_gather(const uint8_t* __restrict__ srcA, const double* __restrict__ fA, int na,
                         const uint8_t* __restrict__ srcB, const double* __restrict__ fB,
                         const int32_t* __restrict__ perm, uint8_t* __restrict__ dst,
                         double* __restrict__ fdst, int N) {
  gather_row(srcA, fA, na, srcB, fB, perm, dst, fdst, N, blockIdx.y);
}

// One population merge of every problem of a batch (the three kernels above,
// problem q in blockIdx.y, blockIdx.x and blockIdx.z respectively).
struct HeaMergeDesc {
  const uint8_t* A;
  const double* fa;
  const uint8_t* B;
  const double* fb;
  int32_t* perm;
  uint32_t* tie;
  uint8_t* dst;
  double* fdst;
  int32_t na, nb, keep, N;
};

__global__ void k_rank_batch(const HeaMergeDesc* __restrict__ md) {
  const HeaMergeDesc m = md[blockIdx.y];
  rank_one(m.fa, m.na, m.fb, m.nb, m.keep, m.perm, m.tie, blockIdx.x * blockDim.x + threadIdx.x);
}

__global__ __launch_bounds__(256) void k_tie_sort_batch(const HeaMergeDesc* __restrict__ md,
                                                        int lds) {
  extern __shared__ __align__(16) unsigned char smem[];
  const HeaMergeDesc m = md[blockIdx.x];
  tie_sort_one(m.fa, m.na, m.fb, m.nb, m.keep, m.perm, m.tie, lds, smem);
}

__global__ void k_gather_batch(const HeaMergeDesc* __restrict__ md) {
  const HeaMergeDesc m = md[blockIdx.z];
  gather_row(m.A, m.fa, m.na, m.B, m.fb, m.perm, m.dst, m.fdst, m.N, blockIdx.y);
}

__global__ void k_iter_add(uint32_t* it, uint32_t v) { *it += v; }
__global__ void k_set_u32(uint32_t* p, uint32_t v) { *p = v; }
__global__ void k_set_u64(unsigned long long* p, unsigned long long v) { *p = v; }
__global__ void k_clock(unsigned long long* t) { *t = wall_clock64(); }

// ------------------------------------------------------------ host objects --

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  hipError_t alloc(size_t count) {
    release();
    n = count;
    if (!count) return hipSuccess;
    return hipMalloc(reinterpret_cast<void**>(&p), sizeof(T) * count);
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~DevBuf() { release(); }
};

// Guided-proposal settings of a local-search launch (gain == nullptr: plain).
struct GuideCfg {
  unsigned long long* hist = nullptr;  // [L][N] per-site neighbour species counts
  int cand = 0;
  double tau = 0.0;
  int multi = 1;
  int32_t* dbg_pairs = nullptr;  // test capture of the first candidate phase
  float* dbg_pred = nullptr;
};

}  // namespace

struct hea_problem_s {
  int device = 0;
  HeaConst hc;
  bool force_hbm_species = false;  // PYHEA_AMD_SPECIES_IN_HBM=1 (tests)
  bool walk_gather = false;        // PYHEA_AMD_WALK_GATHER=1: no neighbour records in the walk
  int guide_blk = 4;               // guided candidates from blocks of 2^b sites (16; 0: uniform);
                                   // PYHEA_AMD_GUIDE_BLOCK=b sets b
  int64_t max_lds = 0;
  DevBuf<HeaConst> d_const;
  DevBuf<int32_t> d_nbr;
  // scratch for the host-buffer entry points (grown on demand)
  DevBuf<uint8_t> d_lat, d_lat2;
  DevBuf<double> d_fit, d_fit2;
  DevBuf<int32_t> d_counts, d_perm;
  DevBuf<int64_t> d_stats;
  DevBuf<uint32_t> d_iter;
  DevBuf<uint32_t> d_tie;  // tie flag of the population sort (zero between calls)
  DevBuf<unsigned long long> d_rec;  // neighbour records of the walk (species in HBM)
  // relaxation: per-workgroup pair-count partials and the slot-major copy
  // of the neighbour table
  DevBuf<double> d_relax_part;
  DevBuf<int32_t> d_relax_nbrT;
};

struct hea_session_s {
  hea_problem p = nullptr;
  int L = 0, task = 0, depth = 0;
  double thr = 0.0;
  uint64_t seed = 0;
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  size_t lds = 0;
  bool sp_lds = true;
  DevBuf<uint8_t> A, B, C;
  DevBuf<double> fA, fB, fC;
  DevBuf<int32_t> counts, perm;
  DevBuf<int64_t> stats;
  DevBuf<uint32_t> iter;
  DevBuf<uint32_t> tie;            // tie flag of the population sorts (zero between sorts)
  DevBuf<unsigned long long> clk;  // [0] start, [1] first fitness <= thr
  int iters_done = 0;
  DevBuf<unsigned long long> hist;  // neighbour records [L][N]: guided mode and/or the walk
  GuideCfg guide;
  unsigned long long* rec = nullptr;  // records of the walk (species in HBM), or nullptr
  int tp = 0;             // proposals per pass (0: default)
};

namespace {

int check_problem(const hea_problem_desc* d, std::string& why) {
  if (!d) return why = "null problem", 0;
  if (d->n_sites < 2 || d->n_sites > (1 << 26)) return why = "n_sites out of range [2, 2^26]", 0;
  if (d->n_species < 2 || d->n_species > HEA_MAX_K)
    return why = "n_species out of range [2, 8]", 0;
  if (d->n_shells < 1 || d->n_shells > HEA_MAX_S) return why = "n_shells out of range [1, 3]", 0;
  if (!d->nbr || !d->counts || !d->weight || !d->target) return why = "null array", 0;
  for (int s = 0; s < d->n_shells; ++s)
    if (d->shell_size[s] < 1 || d->shell_size[s] > HEA_MAX_Z)
      return why = "shell_size out of range [1, 31]", 0;
  int64_t tot = 0;
  for (int a = 0; a < d->n_species; ++a) {
    if (d->counts[a] < 1) return why = "every species needs at least one atom", 0;
    tot += d->counts[a];
  }
  if (tot != d->n_sites) return why = "species counts do not sum to n_sites", 0;
  for (int s = 0; s < d->n_shells; ++s)
    if (!std::isfinite(d->weight[s]) || d->weight[s] < 0.0)
      return why = "weights must be finite and >= 0", 0;
  for (int e = 0; e < d->n_shells * d->n_species * d->n_species; ++e)
    if (!std::isfinite(d->target[e])) return why = "target must be finite", 0;
  return 1;
}

// Segments of R(k) = k-fold sequential sum of delta, verified at every k.
int build_segments(double delta, int64_t kmax, int64_t* k0, double* v0, double* inc, int& nseg) {
  nseg = 0;
  double r = 0.0;        // R(k)
  double rn = delta;     // R(k+1)
  int64_t sk = 0;
  double sv = 0.0, si = rn - r;
  k0[0] = 0;
  v0[0] = 0.0;
  inc[0] = si;
  nseg = 1;
  for (int64_t k = 1; k <= kmax; ++k) {
    r = rn;
    rn = r + delta;
    const double pred = std::fma((double)(k - sk), si, sv);
    if (pred != r) {
      if (nseg == HEA_MAX_SEG) return 0;
      sk = k;
      sv = r;
      si = rn - r;
      k0[nseg] = sk;
      v0[nseg] = sv;
      inc[nseg] = si;
      ++nseg;
    }
  }
  return 1;
}

int ensure_scratch(hea_problem p, int L) {
  const HeaConst& c = p->hc;
  const size_t lat = (size_t)L * c.N;
  if (p->d_lat.n < lat) {
    HEA_HIP(p->d_lat.alloc(lat));
    HEA_HIP(p->d_lat2.alloc(lat));
  }
  if (p->d_fit.n < (size_t)L) {
    HEA_HIP(p->d_fit.alloc(L));
    HEA_HIP(p->d_fit2.alloc(L));
    HEA_HIP(p->d_stats.alloc((size_t)4 * L));
    HEA_HIP(p->d_perm.alloc((size_t)2 * L));
  }
  const size_t cn = (size_t)L * c.S * c.K * c.K;
  if (p->d_counts.n < cn) HEA_HIP(p->d_counts.alloc(cn));
  if (!p->d_iter.p) HEA_HIP(p->d_iter.alloc(1));
  if (!p->d_tie.p) {
    HEA_HIP(p->d_tie.alloc(1));
    HEA_HIP(hipMemset(p->d_tie.p, 0, sizeof(uint32_t)));
  }
  return HEA_OK;
}

int launch_recount(hea_problem p, const uint8_t* lat, int32_t* counts, int L, hipStream_t st) {
  const HeaConst& c = p->hc;
  const size_t lds = recount_lds_bytes(c.N);
  if ((int64_t)lds <= p->max_lds) {  // one workgroup per lattice, species in LDS
    hipLaunchKernelGGL(k_recount_lds, dim3(L), dim3(kRecountLdsThreads), lds, st, p->d_const.p,
                       p->d_nbr.p, lat, counts);
    HEA_HIP(hipGetLastError());
    return HEA_OK;
  }
  const size_t cn = (size_t)L * c.S * c.K * c.K;
  hipLaunchKernelGGL(k_zero_i32, dim3((unsigned)std::min<size_t>((cn + 255) / 256, 4096)),
                     dim3(256), 0, st, counts, cn);
  dim3 g((c.N + kRecountSites - 1) / kRecountSites, L);
  hipLaunchKernelGGL(k_recount, g, dim3(kRecountThreads), 0, st, p->d_const.p, p->d_nbr.p, lat,
                     counts);
  HEA_HIP(hipGetLastError());
  return HEA_OK;
}

int launch_fitness(hea_problem p, const int32_t* counts, double* fit, int L, hipStream_t st) {
  hipLaunchKernelGGL(k_fitness, dim3((L + 63) / 64), dim3(64), 0, st, p->d_const.p, counts, fit,
                     L);
  HEA_HIP(hipGetLastError());
  return HEA_OK;
}

// dst = the reference's order of (fa ++ fb), first `keep` rows.  perm holds
// at least na + nb entries; *tie is zero between calls.
int launch_merge(const uint8_t* A, const double* fa, int na, const uint8_t* B, const double* fb,
                 int nb, int keep, int32_t* perm, uint32_t* tie, uint8_t* dst, double* fdst,
                 int N, hipStream_t st) {
  const int n = na + nb;
  hipLaunchKernelGGL(k_rank, dim3((n + 255) / 256), dim3(256), 0, st, fa, na, fb, nb, keep, perm,
                     tie);
  const size_t tl = (size_t)n * (sizeof(double) + sizeof(int));
  const int lds = tl <= 65536 ? (int)tl : 0;
  hipLaunchKernelGGL(k_tie_sort, dim3(1), dim3(256), (size_t)lds, st, fa, na, fb, nb, keep, perm,
                     tie, lds);
  const int per = (N % 16 == 0) ? N / 16 : N;
  const unsigned gx = (unsigned)std::max(1, std::min(64, (per + 255) / 256));
  hipLaunchKernelGGL(k_gather, dim3(gx, keep), dim3(256), 0, st, A, fa, na, B, fb, perm, dst,
                     fdst, N);
  HEA_HIP(hipGetLastError());
  return HEA_OK;
}

// launch_merge for P problems at once: md[q] is problem q's merge; every
// problem merges na + nb = n rows into keep, its lattices at most maxN sites.
int launch_merge_batch(const HeaMergeDesc* md, int P, int n, int keep, int maxN, hipStream_t st) {
  hipLaunchKernelGGL(k_rank_batch, dim3((n + 255) / 256, P), dim3(256), 0, st, md);
  const size_t tl = (size_t)n * (sizeof(double) + sizeof(int));
  const int lds = tl <= 65536 ? (int)tl : 0;
  hipLaunchKernelGGL(k_tie_sort_batch, dim3(P), dim3(256), (size_t)lds, st, md, lds);
  const int per = (maxN % 16 == 0) ? maxN / 16 : maxN;
  const unsigned gx = (unsigned)std::max(1, std::min(64, (per + 255) / 256));
  hipLaunchKernelGGL(k_gather_batch, dim3(gx, keep, P), dim3(256), 0, st, md);
  HEA_HIP(hipGetLastError());
  return HEA_OK;
}

// Problem scalars of a local-search launch.
void fill_problem_args(hea_problem p, HeaMcArgs& a) {
  const HeaConst& c = p->hc;
  a.pc = p->d_const.p;
  a.nbr = p->d_nbr.p;
  a.N = c.N;
  a.K = c.K;
  a.S = c.S;
  a.stride = c.stride;
  for (int s = 0; s < HEA_MAX_S; ++s) {
    a.Z[s] = c.Z[s];
    a.off[s] = c.off[s];
    a.delta[s] = c.delta[s];
    a.weight[s] = c.weight[s];
  }
  a.g_blk = p->guide_blk;
}

int launch_mc(hea_problem p, uint8_t* lat, double* fit, const int32_t* counts, int64_t* stats,
              const uint32_t* iter_ptr, unsigned long long* first_hit, int L, int task,
              int depth, uint64_t seed, uint32_t phase, double thr, size_t lds, bool sp_lds,
              hipStream_t st, const GuideCfg& gd = GuideCfg(), int tp = 0,
              unsigned long long* rec = nullptr) {
  HeaMcArgs a;
  a.pc = p->d_const.p;
  a.nbr = p->d_nbr.p;
  a.lattices = lat;
  a.fitness = fit;
  a.counts = counts;
  a.stats = stats;
  a.iter_ptr = iter_ptr;
  a.first_hit = first_hit;
  a.seed = seed;
  a.L = L;
  a.task = task;
  a.depth = depth;
  a.tp = tp > 0 ? tp : hea_mc_tp(task, p->hc.S);
  a.phase = phase;
  a.thr = thr;
  a.hist = gd.hist;
  a.g_cand = gd.cand;
  a.g_tau = gd.tau;
  a.g_multi = gd.multi;
  a.g_blk = p->guide_blk;
  a.dbg_pairs = gd.dbg_pairs;
  a.dbg_pred = gd.dbg_pred;
  a.rec = rec;
  const HeaConst& c = p->hc;
  a.N = c.N;
  a.K = c.K;
  a.S = c.S;
  a.stride = c.stride;
  for (int s = 0; s < HEA_MAX_S; ++s) {
    a.Z[s] = c.Z[s];
    a.off[s] = c.off[s];
    a.delta[s] = c.delta[s];
    a.weight[s] = c.weight[s];
  }
  a.dw = hea_mc_draw_lanes(task, c.conc, c.K, hea_mc_split(a, p->hc.self_nbr != 0));
  HEA_HIP(hea_launch_local_mc(a, p->hc.K, sp_lds, p->hc.self_nbr != 0, lds, st));
  return HEA_OK;
}

int check_run_args(hea_problem p, int L, int task, int depth) {
  if (!p) return fail(HEA_EINVAL, "null problem handle");
  if (L < 1) return fail(HEA_EINVAL, "solutions must be >= 1");
  if (task < 1 || task > HEA_MAX_TASK)
    return fail(HEA_EINVAL, "parallel_task must be in [1, 2048]");
  if (depth < 0) return fail(HEA_EINVAL, "converge_depth must be >= 0");
  return HEA_OK;
}

// Whether the walk of a launch with this species placement uses neighbour
// records: species read from HBM, records that fit one word, no override.
bool use_records(hea_problem p, bool sp_lds) {
  return !sp_lds && !p->walk_gather &&
         hea_records_fit(p->hc.K, p->hc.S, p->hc.Z, p->hc.self_nbr != 0);
}

// LDS bytes of a local-search launch; species live in LDS when they fit.
int mc_lds(hea_problem p, int task, size_t* lds, bool* sp_lds, int tp = 0) {
  const bool self = p->hc.self_nbr != 0;
  const int km = hea_mc_kmax(p->hc.K, self);
  if (tp <= 0) tp = hea_mc_tp(task, p->hc.S);
  *sp_lds = !p->force_hbm_species;
  *lds = hea_mc_lds_bytes(km, task, tp, p->hc.N, *sp_lds);
  if ((int64_t)*lds > p->max_lds) {
    *sp_lds = false;  // species read from HBM
    *lds = hea_mc_lds_bytes(km, task, tp, p->hc.N, false);
    if ((int64_t)*lds > p->max_lds) return fail(HEA_EUNSUPPORTED, "parallel_task too large for LDS");
  }
  HEA_HIP(hea_prepare_local_mc(p->hc.K, *sp_lds, self, *lds));
  return HEA_OK;
}

void sum_stats(const std::vector<int64_t>& st, int L, hea_stats* out) {
  hea_stats s = {0, 0, 0, 0, -1.0};
  for (int l = 0; l < L; ++l) {
    s.rounds += st[4 * l + 0];
    s.accepts += st[4 * l + 1];
    s.evals += st[4 * l + 2];
    s.draws += st[4 * l + 3];
  }
  *out = s;
}

}  // namespace

extern "C" {

int hea_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

const char* hea_last_error(void) { return g_err.c_str(); }

int hea_problem_create(const hea_problem_desc* d, int device, hea_problem* out) {
  if (!out) return fail(HEA_EINVAL, "null output handle");
  *out = nullptr;
  std::string why;
  if (!check_problem(d, why)) return fail(HEA_EINVAL, why);
  const int N = d->n_sites, K = d->n_species, S = d->n_shells;
  int sumZ = 0;
  for (int s = 0; s < S; ++s) sumZ += d->shell_size[s];
  for (size_t e = 0; e < (size_t)N * sumZ; ++e)
    if (d->nbr[e] < 0 || d->nbr[e] >= N) return fail(HEA_EINVAL, "neighbour index out of range");
  // The neighbour relation must be symmetric with multiplicity so that the
  // pair counts, and hence gamma, are symmetric.
  int self = 0;
  {
    int o = 0;
    for (int s = 0; s < S; ++s) {
      const int z = d->shell_size[s];
      for (int i = 0; i < N; ++i) {
        const int32_t* li = d->nbr + (size_t)i * sumZ + o;
        for (int k = 0; k < z; ++k) {
          const int j = li[k];
          if (j == i) self = 1;
          const int32_t* lj = d->nbr + (size_t)j * sumZ + o;
          int mij = 0, mji = 0;
          for (int q = 0; q < z; ++q) {
            mij += li[q] == j;
            mji += lj[q] == i;
          }
          if (mij != mji)
            return fail(HEA_EUNSUPPORTED, "neighbour relation is not symmetric (shell " +
                                              std::to_string(s) + ")");
        }
      }
      o += z;
    }
  }
  int dev_count = hea_device_count();
  if (dev_count <= 0) return fail(HEA_ENODEV, "no GPU visible to the HIP runtime");
  if (device < 0 || device >= dev_count) return fail(HEA_EINVAL, "device index out of range");

  hea_problem p = new hea_problem_s();
  p->device = device;
  {
    const char* v = getenv("PYHEA_AMD_SPECIES_IN_HBM");
    p->force_hbm_species = v && v[0] == '1';
    const char* w = getenv("PYHEA_AMD_WALK_GATHER");
    p->walk_gather = w && w[0] == '1';
    const char* g = getenv("PYHEA_AMD_GUIDE_BLOCK");
    if (g && g[0]) p->guide_blk = std::max(0, std::min(6, atoi(g)));
  }
  HeaConst& c = p->hc;
  memset(&c, 0, sizeof(c));
  c.N = N;
  c.K = K;
  c.S = S;
  int off = 0;
  for (int s = 0; s < S; ++s) {
    c.Z[s] = d->shell_size[s];
    c.off[s] = off;
    off += (c.Z[s] + 3) & ~3;
  }
  c.stride = off;
  c.self_nbr = self;
  c.prefix[0] = 0;
  for (int a = 0; a < K; ++a) {
    c.conc[a] = static_cast<double>(d->counts[a]) / static_cast<double>(N);
    c.prefix[a + 1] = c.prefix[a] + d->counts[a];
  }
  for (int i = 0; i < K; ++i)
    for (int j = 0; j < K; ++j) c.dd[i][j] = c.conc[i] * c.conc[j];
  for (int s = 0; s < S; ++s) {
    c.weight[s] = d->weight[s];
    c.delta[s] = 1.0 / (double)(c.Z[s] * N);
    for (int i = 0; i < K; ++i)
      for (int j = 0; j < K; ++j) c.target[s][i][j] = d->target[(size_t)s * K * K + i * K + j];
    if (!build_segments(c.delta[s], (int64_t)c.Z[s] * N, c.seg_k0[s], c.seg_v0[s], c.seg_inc[s],
                        c.nseg[s])) {
      delete p;
      return fail(HEA_EUNSUPPORTED, "gamma sum has too many binade segments");
    }
  }
  // 16 entries of slack after the last row: the walk of a split list may load
  // (and mask) up to three 16-byte chunks past its part
  std::vector<int32_t> padded((size_t)N * c.stride + 16, -1);
  {
    int o = 0;
    for (int s = 0; s < S; ++s) {
      for (int i = 0; i < N; ++i)
        memcpy(&padded[(size_t)i * c.stride + c.off[s]], d->nbr + (size_t)i * sumZ + o,
               sizeof(int32_t) * c.Z[s]);
      o += c.Z[s];
    }
  }
  int rc = HEA_OK;
  do {
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) { rc = fail(HEA_EHIP, hipGetErrorString(e)); break; }
    if ((rc = resolve_scratch(device)) != HEA_OK) break;
    int lds_max = 0;
    e = hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, device);
    if (e != hipSuccess) { rc = fail(HEA_EHIP, hipGetErrorString(e)); break; }
    p->max_lds = lds_max;
    e = hipFuncSetAttribute(reinterpret_cast<const void*>(k_recount_lds),
                            hipFuncAttributeMaxDynamicSharedMemorySize, lds_max);
    if (e == hipSuccess)
      e = hipFuncSetAttribute(reinterpret_cast<const void*>(k_recount_lds_batch),
                              hipFuncAttributeMaxDynamicSharedMemorySize, lds_max);
    if (e != hipSuccess) { rc = fail(HEA_EHIP, hipGetErrorString(e)); break; }
    if (p->d_const.alloc(1) != hipSuccess || p->d_nbr.alloc(padded.size()) != hipSuccess) {
      rc = fail(HEA_ENOMEM, "device allocation failed");
      break;
    }
    e = hipMemcpy(p->d_const.p, &c, sizeof(c), hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = hipMemcpy(p->d_nbr.p, padded.data(), sizeof(int32_t) * padded.size(),
                    hipMemcpyHostToDevice);
    if (e != hipSuccess) { rc = fail(HEA_EHIP, hipGetErrorString(e)); break; }
  } while (0);
  if (rc != HEA_OK) {
    delete p;
    return rc;
  }
  *out = p;
  return HEA_OK;
}

int hea_problem_destroy(hea_problem p) {
  if (!p) return HEA_OK;
  (void)hipSetDevice(p->device);
  delete p;
  return HEA_OK;
}

int hea_fitness(hea_problem p, int32_t L, const uint8_t* lattices, double* out) {
  if (!p || L < 1 || !lattices || !out) return fail(HEA_EINVAL, "bad arguments");
  HEA_HIP(hipSetDevice(p->device));
  int rc = ensure_scratch(p, L);
  if (rc) return rc;
  const size_t n = (size_t)L * p->hc.N;
  HEA_HIP(hipMemcpy(p->d_lat.p, lattices, n, hipMemcpyHostToDevice));
  if ((rc = launch_recount(p, p->d_lat.p, p->d_counts.p, L, nullptr))) return rc;
  if ((rc = launch_fitness(p, p->d_counts.p, p->d_fit.p, L, nullptr))) return rc;
  HEA_HIP(hipMemcpy(out, p->d_fit.p, sizeof(double) * L, hipMemcpyDeviceToHost));
  return HEA_OK;
}

int hea_pair_counts(hea_problem p, int32_t L, const uint8_t* lattices, int32_t* out) {
  if (!p || L < 1 || !lattices || !out) return fail(HEA_EINVAL, "bad arguments");
  HEA_HIP(hipSetDevice(p->device));
  int rc = ensure_scratch(p, L);
  if (rc) return rc;
  HEA_HIP(hipMemcpy(p->d_lat.p, lattices, (size_t)L * p->hc.N, hipMemcpyHostToDevice));
  if ((rc = launch_recount(p, p->d_lat.p, p->d_counts.p, L, nullptr))) return rc;
  const size_t cn = (size_t)L * p->hc.S * p->hc.K * p->hc.K;
  HEA_HIP(hipMemcpy(out, p->d_counts.p, sizeof(int32_t) * cn, hipMemcpyDeviceToHost));
  return HEA_OK;
}

int hea_generate(hea_problem p, int32_t L, uint64_t seed, uint32_t iter, uint32_t phase,
                 uint8_t* out_lat, double* out_fit) {
  if (!p || L < 1 || !out_lat) return fail(HEA_EINVAL, "bad arguments");
  HEA_HIP(hipSetDevice(p->device));
  int rc = ensure_scratch(p, L);
  if (rc) return rc;
  hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, nullptr, p->d_iter.p, iter);
  dim3 g((p->hc.N + 255) / 256, L);
  hipLaunchKernelGGL(k_permute, g, dim3(256), 0, nullptr, p->d_const.p, p->d_lat.p, seed,
                     p->d_iter.p, phase, 0);
  HEA_HIP(hipGetLastError());
  if (out_fit) {
    if ((rc = launch_recount(p, p->d_lat.p, p->d_counts.p, L, nullptr))) return rc;
    if ((rc = launch_fitness(p, p->d_counts.p, p->d_fit.p, L, nullptr))) return rc;
    HEA_HIP(hipMemcpy(out_fit, p->d_fit.p, sizeof(double) * L, hipMemcpyDeviceToHost));
  }
  HEA_HIP(hipMemcpy(out_lat, p->d_lat.p, (size_t)L * p->hc.N, hipMemcpyDeviceToHost));
  return HEA_OK;
}

// Device work of one local search (no sort): iteration word, zeroed
// counters, pair-count recount, search.  iter_w [1], stats [4L] and counts
// [L][S][K][K] are the caller's workspace.  With `records` the walk of
// lattices whose species stay in HBM uses neighbour records kept in the
// problem's scratch (allocated here); without, nothing is allocated.
static int local_mc_enqueue(hea_problem p, int L, uint8_t* lat, double* fit, int task, int depth,
                            uint64_t seed, uint32_t iter, uint32_t phase, uint32_t* iter_w,
                            int64_t* stats, int32_t* counts, hipStream_t st, bool records) {
  size_t lds;
  bool sp_lds;
  int rc = mc_lds(p, task, &lds, &sp_lds);
  if (rc) return rc;
  unsigned long long* rec = nullptr;
  if (records && use_records(p, sp_lds)) {
    const size_t n = (size_t)L * p->hc.N;
    if (p->d_rec.n < n) HEA_HIP(p->d_rec.alloc(n));
    rec = p->d_rec.p;
  }
  hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, st, iter_w, iter);
  hipLaunchKernelGGL(k_zero_i32, dim3(64), dim3(256), 0, st, reinterpret_cast<int32_t*>(stats),
                     (size_t)8 * L);
  if ((rc = launch_recount(p, lat, counts, L, st))) return rc;
  return launch_mc(p, lat, fit, counts, stats, iter_w, nullptr, L, task, depth, seed, phase, -1.0,
                   lds, sp_lds, st, GuideCfg(), 0, rec);
}

static int local_mc_dev(hea_problem p, int L, uint8_t* lat, double* fit, int task, int depth,
                        uint64_t seed, uint32_t iter, uint32_t phase, int sort, hipStream_t st) {
  int rc = ensure_scratch(p, L);
  if (rc) return rc;
  if ((rc = local_mc_enqueue(p, L, lat, fit, task, depth, seed, iter, phase, p->d_iter.p,
                             p->d_stats.p, p->d_counts.p, st, true)))
    return rc;
  if (sort) {
    if ((rc = launch_merge(lat, fit, L, lat, fit, 0, L, p->d_perm.p, p->d_tie.p, p->d_lat2.p,
                           p->d_fit2.p, p->hc.N, st)))
      return rc;
    HEA_HIP(hipMemcpyAsync(lat, p->d_lat2.p, (size_t)L * p->hc.N, hipMemcpyDeviceToDevice, st));
    HEA_HIP(hipMemcpyAsync(fit, p->d_fit2.p, sizeof(double) * L, hipMemcpyDeviceToDevice, st));
  }
  return HEA_OK;
}

int hea_local_mc(hea_problem p, int32_t L, uint8_t* lattices, double* fitness, int32_t task,
                 int32_t depth, uint64_t seed, uint32_t iter, uint32_t phase, int32_t sort,
                 hea_stats* stats) {
  int rc = check_run_args(p, L, task, depth);
  if (rc) return rc;
  if (!lattices || !fitness) return fail(HEA_EINVAL, "null population");
  HEA_HIP(hipSetDevice(p->device));
  const size_t n = (size_t)L * p->hc.N;
  DevBuf<uint8_t> lat;
  DevBuf<double> fit;
  HEA_HIP(lat.alloc(n));
  HEA_HIP(fit.alloc(L));
  HEA_HIP(hipMemcpy(lat.p, lattices, n, hipMemcpyHostToDevice));
  HEA_HIP(hipMemcpy(fit.p, fitness, sizeof(double) * L, hipMemcpyHostToDevice));
  if ((rc = local_mc_dev(p, L, lat.p, fit.p, task, depth, seed, iter, phase, sort, nullptr)))
    return rc;
  HEA_HIP(hipDeviceSynchronize());
  HEA_HIP(hipMemcpy(lattices, lat.p, n, hipMemcpyDeviceToHost));
  HEA_HIP(hipMemcpy(fitness, fit.p, sizeof(double) * L, hipMemcpyDeviceToHost));
  if (stats) {
    std::vector<int64_t> st((size_t)4 * L);
    HEA_HIP(hipMemcpy(st.data(), p->d_stats.p, sizeof(int64_t) * st.size(),
                      hipMemcpyDeviceToHost));
    sum_stats(st, L, stats);
  }
  return HEA_OK;
}

int hea_local_mc_device(hea_problem p, int32_t L, void* d_lattices, void* d_fitness, int32_t task,
                        int32_t depth, uint64_t seed, uint32_t iter, uint32_t phase,
                        void* stream) {
  int rc = check_run_args(p, L, task, depth);
  if (rc) return rc;
  if (!d_lattices || !d_fitness) return fail(HEA_EINVAL, "null population");
  HEA_HIP(hipSetDevice(p->device));
  return local_mc_dev(p, L, static_cast<uint8_t*>(d_lattices), static_cast<double*>(d_fitness),
                      task, depth, seed, iter, phase, 0, static_cast<hipStream_t>(stream));
}

int hea_local_mc_device_noalloc(hea_problem p, int32_t L, void* d_lattices, void* d_fitness,
                                int32_t task, int32_t depth, uint64_t seed, uint32_t iter,
                                uint32_t phase, void* stream) {
  int rc = check_run_args(p, L, task, depth);
  if (rc) return rc;
  if (!d_lattices || !d_fitness) return fail(HEA_EINVAL, "null population");
  // workspace: [iteration word, padded to 256 B][stats int64 4L, padded to
  // 256 B][pair counts int32 L*S*K*K]
  const size_t stats_b = ((size_t)32 * L + 255) & ~(size_t)255;
  const size_t counts_b = (size_t)4 * L * p->hc.S * p->hc.K * p->hc.K;
  const size_t need = 256 + stats_b + counts_b;
  if (need > HEA_NOALLOC_SCRATCH_BYTES)
    return fail(HEA_EINVAL, "solutions too large for the fixed workspace: needs " +
                                std::to_string(need) + " bytes, HEA_NOALLOC_SCRATCH_BYTES = " +
                                std::to_string((size_t)HEA_NOALLOC_SCRATCH_BYTES));
  HEA_HIP(hipSetDevice(p->device));
  unsigned char* ws = scratch_of(p->device);
  if (!ws) return fail(HEA_EHIP, "workspace address of this device was not resolved");
  return local_mc_enqueue(p, L, static_cast<uint8_t*>(d_lattices),
                          static_cast<double*>(d_fitness), task, depth, seed, iter, phase,
                          reinterpret_cast<uint32_t*>(ws), reinterpret_cast<int64_t*>(ws + 256),
                          reinterpret_cast<int32_t*>(ws + 256 + stats_b),
                          static_cast<hipStream_t>(stream), false);
}

// ------------------------------------------------------------------ session --

// Neighbour records of a session's walk: used when its species stay in HBM
// (the buffer is guided mode's records when guided, the same size).
static int session_records(hea_session s) {
  s->rec = nullptr;
  if (!use_records(s->p, s->sp_lds)) return HEA_OK;
  const size_t n = (size_t)s->L * s->p->hc.N;
  if (s->hist.n < n) {
    HEA_HIP(s->hist.alloc(n));
    if (s->guide.hist) s->guide.hist = s->hist.p;
  }
  s->rec = s->hist.p;
  return HEA_OK;
}

// Optional measurement hooks of one iteration: events around both local
// searches and stats snapshots before/after them (eager runs only).
struct IterHooks {
  hipEvent_t ev[4];
  int64_t* snap;  // [3][4L]
};

static int session_enqueue_iteration(hea_session s, const IterHooks* hk = nullptr) {
  hea_problem p = s->p;
  const int L = s->L, N = p->hc.N;
  hipStream_t st = s->stream;
  int rc;
  // new population (accelerate.cpp:743)
  dim3 g((N + 255) / 256, L);
  hipLaunchKernelGGL(k_permute, g, dim3(256), 0, st, p->d_const.p, s->B.p, s->seed, s->iter.p,
                     (uint32_t)HEA_PHASE_NEW, 0);
  if ((rc = launch_recount(p, s->B.p, s->counts.p, L, st))) return rc;
  if ((rc = launch_fitness(p, s->counts.p, s->fB.p, L, st))) return rc;
  // local search I + its closing sort (accelerate.cpp:745, 693)
  const size_t sb = sizeof(int64_t) * 4 * L;
  if (hk) {
    HEA_HIP(hipMemcpyAsync(hk->snap, s->stats.p, sb, hipMemcpyDeviceToDevice, st));
    HEA_HIP(hipEventRecord(hk->ev[0], st));
  }
  if ((rc = launch_mc(p, s->B.p, s->fB.p, s->counts.p, s->stats.p, s->iter.p, s->clk.p + 1, L,
                      s->task, s->depth, s->seed, HEA_PHASE_LS1, s->thr, s->lds, s->sp_lds, st,
                      s->guide, s->tp, s->rec)))
    return rc;
  if (hk) {
    HEA_HIP(hipEventRecord(hk->ev[1], st));
    HEA_HIP(hipMemcpyAsync(hk->snap + 4 * L, s->stats.p, sb, hipMemcpyDeviceToDevice, st));
  }
  if ((rc = launch_merge(s->B.p, s->fB.p, L, s->B.p, s->fB.p, 0, L, s->perm.p, s->tie.p, s->C.p,
                         s->fC.p, N, st)))
    return rc;
  // merge old ++ new, keep the best L (accelerate.cpp:747, 305-330)
  if ((rc = launch_merge(s->A.p, s->fA.p, L, s->C.p, s->fC.p, L, L, s->perm.p, s->tie.p, s->B.p,
                         s->fB.p, N, st)))
    return rc;
  // local search II + sort (accelerate.cpp:749, 693)
  if ((rc = launch_recount(p, s->B.p, s->counts.p, L, st))) return rc;
  if (hk) HEA_HIP(hipEventRecord(hk->ev[2], st));
  if ((rc = launch_mc(p, s->B.p, s->fB.p, s->counts.p, s->stats.p, s->iter.p, s->clk.p + 1, L,
                      s->task, s->depth, s->seed, HEA_PHASE_LS2, s->thr, s->lds, s->sp_lds, st,
                      s->guide, s->tp, s->rec)))
    return rc;
  if (hk) {
    HEA_HIP(hipEventRecord(hk->ev[3], st));
    HEA_HIP(hipMemcpyAsync(hk->snap + 8 * L, s->stats.p, sb, hipMemcpyDeviceToDevice, st));
  }
  if ((rc = launch_merge(s->B.p, s->fB.p, L, s->B.p, s->fB.p, 0, L, s->perm.p, s->tie.p, s->A.p,
                         s->fA.p, N, st)))
    return rc;
  hipLaunchKernelGGL(k_iter_add, dim3(1), dim3(1), 0, st, s->iter.p, 1u);
  HEA_HIP(hipGetLastError());
  return HEA_OK;
}

int hea_session_create(hea_problem p, int32_t L, int32_t task, int32_t depth, double thr,
                       uint64_t seed, hea_session* out) {
  if (!out) return fail(HEA_EINVAL, "null output handle");
  *out = nullptr;
  int rc = check_run_args(p, L, task, depth);
  if (rc) return rc;
  HEA_HIP(hipSetDevice(p->device));
  hea_session s = new hea_session_s();
  s->p = p;
  s->L = L;
  s->task = task;
  s->depth = depth;
  s->thr = thr;
  s->seed = seed;
  const size_t n = (size_t)L * p->hc.N;
  const size_t cn = (size_t)L * p->hc.S * p->hc.K * p->hc.K;
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = s->A.alloc(n);
  if (e == hipSuccess) e = s->B.alloc(n);
  if (e == hipSuccess) e = s->C.alloc(n);
  if (e == hipSuccess) e = s->fA.alloc(L);
  if (e == hipSuccess) e = s->fB.alloc(L);
  if (e == hipSuccess) e = s->fC.alloc(L);
  if (e == hipSuccess) e = s->counts.alloc(cn);
  if (e == hipSuccess) e = s->perm.alloc((size_t)2 * L);
  if (e == hipSuccess) e = s->stats.alloc((size_t)4 * L);
  if (e == hipSuccess) e = s->iter.alloc(1);
  if (e == hipSuccess) e = s->clk.alloc(2);
  if (e == hipSuccess) e = s->tie.alloc(1);
  if (e == hipSuccess) e = hipMemset(s->tie.p, 0, sizeof(uint32_t));
  if (e != hipSuccess) {
    delete s;
    return fail(HEA_ENOMEM, std::string("session allocation: ") + hipGetErrorString(e));
  }
  if ((rc = mc_lds(p, task, &s->lds, &s->sp_lds)) || (rc = session_records(s))) {
    delete s;
    return rc;
  }
  e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete s;
    return fail(HEA_EHIP, hipGetErrorString(e));
  }
  *out = s;
  return HEA_OK;
}

static int valid_tp(int tp, int S) {
  return tp >= 64 && tp <= HEA_MC_MAX_TP && tp % 64 == 0 && S * tp <= HEA_MC_MAX_THREADS;
}

int hea_session_tile(hea_session s, int32_t tp) {
  if (!s || !valid_tp(tp, s->p->hc.S))
    return fail(HEA_EINVAL, "tile must be 64, 128, 192 or 256");
  HEA_HIP(hipSetDevice(s->p->device));
  int rc;
  size_t lds;
  bool sp;
  if ((rc = mc_lds(s->p, s->task, &lds, &sp, tp))) return rc;
  s->tp = tp;
  s->lds = lds;
  s->sp_lds = sp;
  if ((rc = session_records(s))) return rc;
  if (s->exec) {
    (void)hipGraphExecDestroy(s->exec);
    (void)hipGraphDestroy(s->graph);
    s->exec = nullptr;
    s->graph = nullptr;
  }
  return HEA_OK;
}

static int check_guide(hea_problem p, int candidates, double tau, int accepts) {
  if (candidates < 1 || !(tau > 0.0) || accepts < 1)
    return fail(HEA_EINVAL, "guide needs candidates >= 1, tau > 0, accepts >= 1");
  if (p->hc.K > 6) return fail(HEA_EUNSUPPORTED, "guided proposals support up to 6 species");
  if (p->hc.self_nbr) return fail(HEA_EUNSUPPORTED, "guided proposals need cells without self neighbours");
  const HeaHistLayout h = hea_hist_layout(p->hc.S, p->hc.Z);
  if ((p->hc.K - 1) * h.W > 64)
    return fail(HEA_EUNSUPPORTED, "guided proposals need (K-1) * sum_s bitwidth(Z_s) <= 64");
  return HEA_OK;
}

int hea_guided_scores(hea_problem p, int32_t L, const uint8_t* lattices, int32_t task,
                      int32_t candidates, uint64_t seed, uint32_t iter, uint32_t phase,
                      int32_t* out_pairs, float* out_pred) {
  int rc = check_run_args(p, L, task, 1);
  if (rc) return rc;
  if ((rc = check_guide(p, candidates, 0.1, 1))) return rc;
  if (!lattices || !out_pairs || !out_pred) return fail(HEA_EINVAL, "null buffer");
  HEA_HIP(hipSetDevice(p->device));
  if ((rc = ensure_scratch(p, L))) return rc;
  const size_t n = (size_t)L * p->hc.N, items = (size_t)L * task * candidates;
  DevBuf<unsigned long long> hist;
  DevBuf<int32_t> pairs;
  DevBuf<float> pred;
  HEA_HIP(hist.alloc(n));
  HEA_HIP(pairs.alloc(2 * items));
  HEA_HIP(pred.alloc(items));
  HEA_HIP(hipMemcpy(p->d_lat.p, lattices, n, hipMemcpyHostToDevice));
  if ((rc = launch_recount(p, p->d_lat.p, p->d_counts.p, L, nullptr))) return rc;
  if ((rc = launch_fitness(p, p->d_counts.p, p->d_fit.p, L, nullptr))) return rc;
  hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, nullptr, p->d_iter.p, iter);
  hipLaunchKernelGGL(k_zero_i32, dim3(64), dim3(256), 0, nullptr,
                     reinterpret_cast<int32_t*>(p->d_stats.p), (size_t)8 * L);
  size_t lds;
  bool sp_lds;
  if ((rc = mc_lds(p, task, &lds, &sp_lds))) return rc;
  GuideCfg g;
  g.hist = hist.p;
  g.cand = candidates;
  g.tau = 0.1;
  g.multi = 1;
  g.dbg_pairs = pairs.p;
  g.dbg_pred = pred.p;
  if ((rc = launch_mc(p, p->d_lat.p, p->d_fit.p, p->d_counts.p, p->d_stats.p, p->d_iter.p,
                      nullptr, L, task, 1, seed, phase, -1.0, lds, sp_lds, nullptr, g)))
    return rc;
  HEA_HIP(hipDeviceSynchronize());
  HEA_HIP(hipMemcpy(out_pairs, pairs.p, sizeof(int32_t) * 2 * items, hipMemcpyDeviceToHost));
  HEA_HIP(hipMemcpy(out_pred, pred.p, sizeof(float) * items, hipMemcpyDeviceToHost));
  return HEA_OK;
}

int hea_session_guide(hea_session s, int32_t candidates, double tau, int32_t accepts) {
  if (!s) return fail(HEA_EINVAL, "null session");
  int rc = check_guide(s->p, candidates, tau, accepts);
  if (rc) return rc;
  HEA_HIP(hipSetDevice(s->p->device));
  const size_t n = (size_t)s->L * s->p->hc.N;
  if (s->hist.n < n) HEA_HIP(s->hist.alloc(n));
  size_t lds;
  bool sp;
  if ((rc = mc_lds(s->p, s->task, &lds, &sp, s->tp))) return rc;
  s->lds = lds;
  s->sp_lds = sp;
  s->guide.hist = s->hist.p;
  s->guide.cand = candidates;
  s->guide.tau = tau;
  s->guide.multi = accepts;
  if ((rc = session_records(s))) return rc;
  if (s->exec) {  // the graph bakes the launch arguments in: capture again
    (void)hipGraphExecDestroy(s->exec);
    (void)hipGraphDestroy(s->graph);
    s->exec = nullptr;
    s->graph = nullptr;
  }
  return HEA_OK;
}

int hea_session_init(hea_session s) {
  if (!s) return fail(HEA_EINVAL, "null session");
  hea_problem p = s->p;
  HEA_HIP(hipSetDevice(p->device));
  const int L = s->L;
  hipStream_t st = s->stream;
  hipLaunchKernelGGL(k_zero_i32, dim3(64), dim3(256), 0, st,
                     reinterpret_cast<int32_t*>(s->stats.p), (size_t)8 * L);
  hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, st, s->iter.p, 0u);
  hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, st, s->clk.p + 1, ~0ull);
  hipLaunchKernelGGL(k_clock, dim3(1), dim3(1), 0, st, s->clk.p);
  // initial population (accelerate.cpp:737)
  dim3 g((p->hc.N + 255) / 256, L);
  hipLaunchKernelGGL(k_permute, g, dim3(256), 0, st, p->d_const.p, s->A.p, s->seed, s->iter.p,
                     (uint32_t)HEA_PHASE_INIT, 0);
  int rc;
  if ((rc = launch_recount(p, s->A.p, s->counts.p, L, st))) return rc;
  if ((rc = launch_fitness(p, s->counts.p, s->fA.p, L, st))) return rc;
  HEA_HIP(hipGetLastError());
  s->iters_done = 0;
  return HEA_OK;
}

int hea_session_step(hea_session s, int32_t n_iter) {
  if (!s || n_iter < 0) return fail(HEA_EINVAL, "bad arguments");
  HEA_HIP(hipSetDevice(s->p->device));
  if (!s->exec) {
    HEA_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    int rc = session_enqueue_iteration(s);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return fail(HEA_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
    s->graph = g;
    HEA_HIP(hipGraphInstantiate(&s->exec, s->graph, nullptr, nullptr, 0));
  }
  for (int i = 0; i < n_iter; ++i) HEA_HIP(hipGraphLaunch(s->exec, s->stream));
  s->iters_done += n_iter;
  return HEA_OK;
}

int hea_session_step_timed(hea_session s, int32_t n_iter, float* out_ms, int64_t* out_evals) {
  if (!s || n_iter < 0 || !out_ms || !out_evals) return fail(HEA_EINVAL, "bad arguments");
  HEA_HIP(hipSetDevice(s->p->device));
  const int L = s->L;
  IterHooks hk;
  for (auto& e : hk.ev) HEA_HIP(hipEventCreate(&e));
  DevBuf<int64_t> snap;
  HEA_HIP(snap.alloc((size_t)12 * L));
  hk.snap = snap.p;
  std::vector<int64_t> h((size_t)12 * L);
  int rc = HEA_OK;
  for (int i = 0; i < n_iter && rc == HEA_OK; ++i) {
    if ((rc = session_enqueue_iteration(s, &hk))) break;
    hipError_t e = hipStreamSynchronize(s->stream);
    if (e == hipSuccess) e = hipEventElapsedTime(&out_ms[2 * i], hk.ev[0], hk.ev[1]);
    if (e == hipSuccess) e = hipEventElapsedTime(&out_ms[2 * i + 1], hk.ev[2], hk.ev[3]);
    if (e == hipSuccess)
      e = hipMemcpy(h.data(), snap.p, sizeof(int64_t) * h.size(), hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
      rc = fail(HEA_EHIP, hipGetErrorString(e));
      break;
    }
    int64_t ev[3] = {0, 0, 0};
    for (int q = 0; q < 3; ++q)
      for (int l = 0; l < L; ++l) ev[q] += h[(size_t)q * 4 * L + 4 * l + 2];
    out_evals[2 * i] = ev[1] - ev[0];
    out_evals[2 * i + 1] = ev[2] - ev[1];
    ++s->iters_done;
  }
  for (auto& e : hk.ev) (void)hipEventDestroy(e);
  return rc;
}

int hea_session_sync(hea_session s) {
  if (!s) return fail(HEA_EINVAL, "null session");
  HEA_HIP(hipSetDevice(s->p->device));
  HEA_HIP(hipStreamSynchronize(s->stream));
  return HEA_OK;
}

int hea_session_best(hea_session s, double* best) {
  if (!s || !best) return fail(HEA_EINVAL, "bad arguments");
  HEA_HIP(hipSetDevice(s->p->device));
  HEA_HIP(hipStreamSynchronize(s->stream));
  HEA_HIP(hipMemcpy(best, s->fA.p, sizeof(double), hipMemcpyDeviceToHost));
  return HEA_OK;
}

int hea_session_read(hea_session s, uint8_t* out_lat, double* out_fit, hea_stats* stats) {
  if (!s) return fail(HEA_EINVAL, "null session");
  hea_problem p = s->p;
  HEA_HIP(hipSetDevice(p->device));
  HEA_HIP(hipStreamSynchronize(s->stream));
  const int L = s->L;
  if (out_lat) HEA_HIP(hipMemcpy(out_lat, s->A.p, (size_t)L * p->hc.N, hipMemcpyDeviceToHost));
  if (out_fit) HEA_HIP(hipMemcpy(out_fit, s->fA.p, sizeof(double) * L, hipMemcpyDeviceToHost));
  if (stats) {
    std::vector<int64_t> st((size_t)4 * L);
    HEA_HIP(hipMemcpy(st.data(), s->stats.p, sizeof(int64_t) * st.size(),
                      hipMemcpyDeviceToHost));
    sum_stats(st, L, stats);
    unsigned long long clk[2];
    HEA_HIP(hipMemcpy(clk, s->clk.p, sizeof(clk), hipMemcpyDeviceToHost));
    int khz = 0;
    HEA_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, p->device));
    if (clk[1] != ~0ull && khz > 0)
      stats->first_hit_s = (double)(clk[1] - clk[0]) / (1000.0 * khz);
  }
  return HEA_OK;
}

int hea_session_destroy(hea_session s) {
  if (!s) return HEA_OK;
  (void)hipSetDevice(s->p->device);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  if (s->exec) (void)hipGraphExecDestroy(s->exec);
  if (s->graph) (void)hipGraphDestroy(s->graph);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return HEA_OK;
}

int hea_hcs_run(hea_problem p, int32_t L, int32_t total_iter, int32_t task, int32_t depth,
                double thr, uint64_t seed, uint8_t* out_lat, double* out_fit, hea_stats* stats) {
  if (total_iter < 0) return fail(HEA_EINVAL, "total_iter must be >= 0");
  hea_session s = nullptr;
  int rc = hea_session_create(p, L, task, depth, thr, seed, &s);
  if (rc) return rc;
  if (!(rc = hea_session_init(s)) && !(rc = hea_session_step(s, total_iter)))
    rc = hea_session_read(s, out_lat, out_fit, stats);
  hea_session_destroy(s);
  return rc;
}

}  // extern "C"

// ------------------------------------------------------------------- batch --
// Many independent problems searched together.  Problem p keeps its own
// population (A: current, B: searched, C: sorted scratch) and draws from seed
// `seed + p`, so its result equals hea_hcs_run(problem p, seed + p).  The
// cheap phases are one launch per step for all problems; both local
// searches run as one persistent launch over all (problem, lattice) items,
// longest first (a batch of one problem launches the per-lattice kernel
// instead, its arguments in the kernarg segment and its table width exact).
// One global iteration is captured as one HIP graph.

struct hea_batch_s {
  std::vector<hea_problem> probs;
  int L = 0, task = 0, depth = 0, tp = 0, S = 0, kmax = 0, threads = 0, grid = 0;
  double thr = 0.0;
  uint64_t seed = 0;
  bool sp_lds = true;
  size_t lds = 0;
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<size_t> lat_off, cnt_off;  // element offsets per problem
  DevBuf<uint8_t> A, B, C;
  DevBuf<double> fA, fB, fC;
  DevBuf<int32_t> counts, perm;
  DevBuf<int64_t> stats;
  DevBuf<uint32_t> iter, queue;   // queue: [HEA_XCDS] next item of each part
  int xoff[HEA_XCDS + 1] = {};     // item parts (HeaMcBatchArgs)
  DevBuf<uint32_t> tie;            // [problems] tie flags of the population sorts
  DevBuf<unsigned long long> clk;  // [0] start, [1 + p] first fitness <= thr of problem p
  DevBuf<HeaMcArgs> margs;         // launch arguments [2 phases][problems] for population B
  std::vector<HeaMcArgs> hmargs;   // host copy of margs: a batch of one problem launches
                                   // k_local_mc with them as kernel arguments
  DevBuf<HeaMergeDesc> merges;     // [3][problems]: B -> C and A ++ C -> B after local
                                   // search I, B -> A after local search II
  DevBuf<HeaPopDesc> pops;         // [problems]: population B, its counts and fitness
  int maxN = 0;
  size_t recount_lds = 0;          // dynamic LDS of the batched recount (0: per problem)
  DevBuf<int2> items;
  DevBuf<unsigned long long> hist;  // guided mode, [sum_p L N_p]
  GuideCfg guide;
};

extern "C" int hea_batch_grid(hea_batch b, int32_t grid);  // autotuner hook (below)
static int batch_records(hea_batch b);

// Pair counts of population B of every problem: one launch when every
// problem's lattices fit in LDS, else one recount per problem.
static int batch_recount(hea_batch b) {
  const int L = b->L, P = (int)b->probs.size();
  if (b->recount_lds) {
    hipLaunchKernelGGL(k_recount_lds_batch, dim3(L, P), dim3(kRecountLdsThreads), b->recount_lds,
                       b->stream, b->pops.p);
    HEA_HIP(hipGetLastError());
    return HEA_OK;
  }
  int rc;
  for (int q = 0; q < P; ++q)
    if ((rc = launch_recount(b->probs[q], b->B.p + b->lat_off[q], b->counts.p + b->cnt_off[q], L,
                             b->stream)))
      return rc;
  return HEA_OK;
}

static int batch_enqueue_iteration(hea_batch b, const IterHooks* hk = nullptr) {
  const int L = b->L, P = (int)b->probs.size();
  const size_t sn = (size_t)4 * L * P;  // stats entries of the whole batch
  hipStream_t st = b->stream;
  int rc;
  const unsigned gx = (unsigned)std::min(16, (b->maxN + 255) / 256);
  hipLaunchKernelGGL(k_permute_batch, dim3(gx, L, P), dim3(256), 0, st, b->pops.p, b->iter.p,
                     (uint32_t)HEA_PHASE_NEW);
  if ((rc = batch_recount(b))) return rc;
  hipLaunchKernelGGL(k_fitness_batch, dim3((L + 63) / 64, P), dim3(64), 0, st, b->pops.p, L);
  HeaMcBatchArgs ba;
  ba.items = b->items.p;
  ba.n_items = P * L;
  for (int x = 0; x <= HEA_XCDS; ++x) ba.xoff[x] = b->xoff[x];
  ba.queue = b->queue.p;
  for (int phase = HEA_PHASE_LS1; phase <= HEA_PHASE_LS2; ++phase) {
    ba.probs = b->margs.p + (size_t)(phase - HEA_PHASE_LS1) * P;
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, st,
                       reinterpret_cast<int32_t*>(b->queue.p), (size_t)HEA_XCDS);
    const int h = phase - HEA_PHASE_LS1;
    if (hk) {
      if (h == 0)
        HEA_HIP(hipMemcpyAsync(hk->snap, b->stats.p, sizeof(int64_t) * sn,
                               hipMemcpyDeviceToDevice, st));
      HEA_HIP(hipEventRecord(hk->ev[2 * h], st));
    }
    if (P == 1)  // one problem: nothing to balance, its arguments go in the kernarg segment
      HEA_HIP(hea_launch_local_mc(b->hmargs[h], b->probs[0]->hc.K, b->sp_lds, false, b->lds, st));
    else
      HEA_HIP(hea_launch_local_mc_batch(ba, b->kmax, b->sp_lds, b->guide.hist != nullptr, b->grid,
                                        b->threads, b->lds, st));
    if (hk) {
      HEA_HIP(hipEventRecord(hk->ev[2 * h + 1], st));
      HEA_HIP(hipMemcpyAsync(hk->snap + (h + 1) * sn, b->stats.p, sizeof(int64_t) * sn,
                             hipMemcpyDeviceToDevice, st));
    }
    if (phase == HEA_PHASE_LS1) {
      if ((rc = launch_merge_batch(b->merges.p, P, L, L, b->maxN, st))) return rc;
      if ((rc = launch_merge_batch(b->merges.p + P, P, 2 * L, L, b->maxN, st))) return rc;
      if ((rc = batch_recount(b))) return rc;
    } else {
      if ((rc = launch_merge_batch(b->merges.p + 2 * P, P, L, L, b->maxN, st))) return rc;
    }
  }
  hipLaunchKernelGGL(k_iter_add, dim3(1), dim3(1), 0, st, b->iter.p, 1u);
  HEA_HIP(hipGetLastError());
  return HEA_OK;
}

extern "C" {

int hea_batch_create(const hea_problem* probs, int32_t n_probs, int32_t L, int32_t task,
                     int32_t depth, double thr, uint64_t seed, hea_batch* out) {
  if (!out) return fail(HEA_EINVAL, "null output handle");
  *out = nullptr;
  if (!probs || n_probs < 1) return fail(HEA_EINVAL, "need at least one problem");
  if (n_probs > 65535)  // the batched merge kernels index problems by a grid dimension
    return fail(HEA_EINVAL, "at most 65535 problems per batch");
  int rc;
  for (int q = 0; q < n_probs; ++q)
    if ((rc = check_run_args(probs[q], L, task, depth))) return rc;
  const int dev = probs[0]->device, S = probs[0]->hc.S;
  int kmax_need = 0, maxN = 0;
  for (int q = 0; q < n_probs; ++q) {
    if (probs[q]->device != dev) return fail(HEA_EINVAL, "batch problems must share a device");
    if (probs[q]->hc.S != S) return fail(HEA_EINVAL, "batch problems must share the shell count");
    if (probs[q]->hc.self_nbr)
      return fail(HEA_EUNSUPPORTED, "batched search needs cells without self neighbours");
    kmax_need = std::max(kmax_need, probs[q]->hc.K);
    maxN = std::max(maxN, probs[q]->hc.N);
  }
  HEA_HIP(hipSetDevice(dev));
  hea_batch b = new hea_batch_s();
  b->probs.assign(probs, probs + n_probs);
  b->L = L;
  b->task = task;
  b->depth = depth;
  b->thr = thr;
  b->seed = seed;
  b->S = S;
  b->kmax = kmax_need <= 5 ? 5 : kmax_need <= 6 ? 6 : 8;  // batch kernel widths
  b->tp = hea_mc_tp(task, S);
  b->threads = 0;  // the largest screen lane count of the problems (set with their args)
  b->sp_lds = !probs[0]->force_hbm_species;
  b->lds = hea_mc_lds_bytes(b->kmax, task, b->tp, maxN, b->sp_lds);
  if ((int64_t)b->lds > probs[0]->max_lds) {
    b->sp_lds = false;
    b->lds = hea_mc_lds_bytes(b->kmax, task, b->tp, maxN, false);
  }
  size_t nl = 0, nc = 0;
  std::vector<double> cost(n_probs);
  for (int q = 0; q < n_probs; ++q) {
    const HeaConst& c = probs[q]->hc;
    b->lat_off.push_back(nl);
    b->cnt_off.push_back(nc);
    nl += (size_t)L * c.N;
    nc += (size_t)L * c.S * c.K * c.K;
    int sz = 0;
    for (int s = 0; s < c.S; ++s) sz += c.Z[s];
    cost[q] = (double)c.N * (sz + c.K * c.K);  // rounds ~ N, round ~ sumZ + K^2
  }
  // one part per XCD: problems, largest first, each to the least loaded
  // part (all of a problem's lattices together, so its neighbour table is
  // read through one L2); within a part, longest items first
  std::vector<int> order(n_probs), part(n_probs);
  for (int q = 0; q < n_probs; ++q) order[q] = q;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost[x] > cost[y]; });
  double load[HEA_XCDS] = {};
  for (int q : order) {
    int x = 0;
    for (int k = 1; k < HEA_XCDS; ++k)
      if (load[k] < load[x]) x = k;
    part[q] = x;
    load[x] += cost[q] * L;
  }
  std::vector<int2> items;
  for (int x = 0; x < HEA_XCDS; ++x) {
    b->xoff[x] = (int)items.size();
    for (int q : order)
      if (part[q] == x)
        for (int l = 0; l < L; ++l) items.push_back(make_int2(q, l));
  }
  b->xoff[HEA_XCDS] = (int)items.size();
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = b->A.alloc(nl);
  if (e == hipSuccess) e = b->B.alloc(nl);
  if (e == hipSuccess) e = b->C.alloc(nl);
  if (e == hipSuccess) e = b->fA.alloc((size_t)n_probs * L);
  if (e == hipSuccess) e = b->fB.alloc((size_t)n_probs * L);
  if (e == hipSuccess) e = b->fC.alloc((size_t)n_probs * L);
  if (e == hipSuccess) e = b->counts.alloc(nc);
  if (e == hipSuccess) e = b->perm.alloc((size_t)n_probs * 2 * L);
  if (e == hipSuccess) e = b->stats.alloc((size_t)n_probs * 4 * L);
  if (e == hipSuccess) e = b->iter.alloc(1);
  if (e == hipSuccess) e = b->queue.alloc(HEA_XCDS);
  if (e == hipSuccess) e = b->clk.alloc((size_t)1 + n_probs);
  if (e == hipSuccess) e = b->margs.alloc((size_t)2 * n_probs);
  if (e == hipSuccess) e = b->merges.alloc((size_t)3 * n_probs);
  if (e == hipSuccess) e = b->pops.alloc((size_t)n_probs);
  if (e == hipSuccess) e = b->items.alloc(items.size());
  if (e == hipSuccess) e = b->tie.alloc((size_t)n_probs);
  if (e == hipSuccess) e = hipMemset(b->tie.p, 0, sizeof(uint32_t) * n_probs);
  if (e != hipSuccess) {
    delete b;
    return fail(HEA_ENOMEM, std::string("batch allocation: ") + hipGetErrorString(e));
  }
  std::vector<HeaMcArgs> ma((size_t)2 * n_probs);
  for (int ph = 0; ph < 2; ++ph)
    for (int q = 0; q < n_probs; ++q) {
      HeaMcArgs& a = ma[(size_t)ph * n_probs + q];
      memset(&a, 0, sizeof(a));
      fill_problem_args(probs[q], a);
      a.lattices = b->B.p + b->lat_off[q];
      a.fitness = b->fB.p + (size_t)q * L;
      a.counts = b->counts.p + b->cnt_off[q];
      a.stats = b->stats.p + (size_t)q * 4 * L;
      a.first_hit = b->clk.p + 1 + q;
      a.iter_ptr = b->iter.p;
      a.L = L;
      a.task = task;
      a.depth = depth;
      a.tp = b->tp;
      const int lanes = hea_mc_split(a, false);
      a.dw = hea_mc_draw_lanes(task, probs[q]->hc.conc, probs[q]->hc.K, lanes);
      b->threads = std::max(b->threads, lanes);
      a.thr = thr;
      a.seed = seed + (uint64_t)q;
      a.phase = (uint32_t)(HEA_PHASE_LS1 + ph);
    }
  b->maxN = maxN;
  std::vector<HeaMergeDesc> md((size_t)3 * n_probs);
  for (int q = 0; q < n_probs; ++q) {
    uint8_t* A = b->A.p + b->lat_off[q];
    uint8_t* B = b->B.p + b->lat_off[q];
    uint8_t* C = b->C.p + b->lat_off[q];
    double* fA = b->fA.p + (size_t)q * L;
    double* fB = b->fB.p + (size_t)q * L;
    double* fC = b->fC.p + (size_t)q * L;
    int32_t* perm = b->perm.p + (size_t)q * 2 * L;
    uint32_t* tie = b->tie.p + q;
    const int N = probs[q]->hc.N;
    md[q] = HeaMergeDesc{B, fB, B, fB, perm, tie, C, fC, L, 0, L, N};
    md[(size_t)n_probs + q] = HeaMergeDesc{A, fA, C, fC, perm, tie, B, fB, L, L, L, N};
    md[(size_t)2 * n_probs + q] = HeaMergeDesc{B, fB, B, fB, perm, tie, A, fA, L, 0, L, N};
  }
  std::vector<HeaPopDesc> pd(n_probs);
  for (int q = 0; q < n_probs; ++q)
    pd[q] = HeaPopDesc{probs[q]->d_const.p, probs[q]->d_nbr.p, b->B.p + b->lat_off[q],
                       b->counts.p + b->cnt_off[q], b->fB.p + (size_t)q * L, seed + (uint64_t)q};
  {
    const size_t lds = recount_lds_bytes(maxN);
    b->recount_lds = (int64_t)lds <= probs[0]->max_lds ? lds : 0;
  }
  e = hipMemcpy(b->margs.p, ma.data(), sizeof(HeaMcArgs) * ma.size(), hipMemcpyHostToDevice);
  b->hmargs = ma;
  if (e == hipSuccess)
    e = hipMemcpy(b->merges.p, md.data(), sizeof(HeaMergeDesc) * md.size(),
                  hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(b->pops.p, pd.data(), sizeof(HeaPopDesc) * pd.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(b->items.p, items.data(), sizeof(int2) * items.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hea_prepare_local_mc_batch(b->kmax, b->sp_lds, b->lds);
  if (e == hipSuccess && n_probs == 1)
    e = hea_prepare_local_mc(probs[0]->hc.K, b->sp_lds, false, b->lds);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete b;
    return fail(HEA_EHIP, hipGetErrorString(e));
  }
  int cus = 0;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  b->grid = std::min((int)items.size(), std::max(1, cus) * 2);
  if ((rc = batch_records(b))) {
    delete b;
    return rc;
  }
  *out = b;
  return HEA_OK;
}

int hea_batch_grid(hea_batch b, int32_t grid) {
  if (!b || grid < 1) return fail(HEA_EINVAL, "bad arguments");
  b->grid = std::min(grid, (int32_t)(b->probs.size() * b->L));
  if (b->exec) {  // the graph bakes the launch shape in: capture again
    (void)hipGraphExecDestroy(b->exec);
    (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
  }
  return HEA_OK;
}

// Rewrites the batch's device launch arguments with `patch` applied.
static int batch_patch_args(hea_batch b, const std::function<void(HeaMcArgs&, int)>& patch) {
  const size_t n = (size_t)2 * b->probs.size();
  std::vector<HeaMcArgs> ma(n);
  HEA_HIP(hipMemcpy(ma.data(), b->margs.p, sizeof(HeaMcArgs) * n, hipMemcpyDeviceToHost));
  for (size_t i = 0; i < n; ++i) patch(ma[i], (int)(i % b->probs.size()));
  HEA_HIP(hipMemcpy(b->margs.p, ma.data(), sizeof(HeaMcArgs) * n, hipMemcpyHostToDevice));
  b->hmargs = ma;
  return HEA_OK;
}

// Neighbour records of the batch's walks: problems whose species stay in HBM
// use their slice of b->hist (guided mode's records when guided).
static int batch_records(hea_batch b) {
  const int P = (int)b->probs.size();
  std::vector<char> use(P);
  std::vector<size_t> off(P);
  size_t n = 0;
  bool any = false;
  for (int q = 0; q < P; ++q) {
    use[q] = use_records(b->probs[q], b->sp_lds);
    any = any || use[q];
    off[q] = n;
    n += (size_t)b->L * b->probs[q]->hc.N;
  }
  if (any && b->hist.n < n) {
    if (b->guide.hist) return fail(HEA_EINVAL, "guided records buffer too small");
    HEA_HIP(b->hist.alloc(n));
  }
  unsigned long long* base = b->hist.p;
  b->threads = 0;  // the screen lanes depend on the walk (hea_mc_split)
  return batch_patch_args(b, [&](HeaMcArgs& a, int q) {
    a.rec = use[q] ? base + off[q] : nullptr;
    const int lanes = hea_mc_split(a, false);
    a.dw = hea_mc_draw_lanes(b->task, b->probs[q]->hc.conc, b->probs[q]->hc.K, lanes);
    b->threads = std::max(b->threads, lanes);
  });
}

int hea_batch_tile(hea_batch b, int32_t tp) {
  if (!b || !valid_tp(tp, b->S))
    return fail(HEA_EINVAL, "tile must be 64, 128, 192 or 256");
  int maxN = 0;
  for (auto q : b->probs) maxN = std::max(maxN, q->hc.N);
  b->tp = tp;
  b->threads = 0;
  b->sp_lds = !b->probs[0]->force_hbm_species;
  b->lds = hea_mc_lds_bytes(b->kmax, b->task, tp, maxN, b->sp_lds);
  if ((int64_t)b->lds > b->probs[0]->max_lds) {
    b->sp_lds = false;
    b->lds = hea_mc_lds_bytes(b->kmax, b->task, tp, maxN, false);
  }
  HEA_HIP(hipSetDevice(b->probs[0]->device));
  HEA_HIP(hea_prepare_local_mc_batch(b->kmax, b->sp_lds, b->lds));
  if (b->probs.size() == 1)
    HEA_HIP(hea_prepare_local_mc(b->probs[0]->hc.K, b->sp_lds, false, b->lds));
  int rc = batch_patch_args(b, [tp, b](HeaMcArgs& a, int q) {
    a.tp = tp;
    const int lanes = hea_mc_split(a, false);
    a.dw = hea_mc_draw_lanes(b->task, b->probs[q]->hc.conc, b->probs[q]->hc.K, lanes);
    b->threads = std::max(b->threads, lanes);
  });
  if (!rc) rc = batch_records(b);
  if (rc) return rc;
  if (b->exec) {
    (void)hipGraphExecDestroy(b->exec);
    (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
  }
  return HEA_OK;
}

int hea_batch_guide(hea_batch b, int32_t candidates, double tau, int32_t accepts) {
  if (!b) return fail(HEA_EINVAL, "null batch");
  int rc;
  for (auto q : b->probs)
    if ((rc = check_guide(q, candidates, tau, accepts))) return rc;
  HEA_HIP(hipSetDevice(b->probs[0]->device));
  const int P = (int)b->probs.size();
  size_t n = 0;
  std::vector<size_t> off(P);
  for (int q = 0; q < P; ++q) {
    off[q] = n;
    n += (size_t)b->L * b->probs[q]->hc.N;
  }
  if (b->hist.n < n) HEA_HIP(b->hist.alloc(n));
  unsigned long long* base = b->hist.p;
  rc = batch_patch_args(b, [&](HeaMcArgs& a, int q) {
    a.hist = base + off[q];
    a.g_cand = candidates;
    a.g_tau = tau;
    a.g_multi = accepts;
  });
  if (rc) return rc;
  b->guide.hist = b->hist.p;
  if ((rc = batch_records(b))) return rc;
  b->guide.cand = candidates;
  b->guide.tau = tau;
  b->guide.multi = accepts;
  if (b->exec) {
    (void)hipGraphExecDestroy(b->exec);
    (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
  }
  return HEA_OK;
}

int hea_batch_init(hea_batch b) {
  if (!b) return fail(HEA_EINVAL, "null batch");
  HEA_HIP(hipSetDevice(b->probs[0]->device));
  const int L = b->L, P = (int)b->probs.size();
  hipStream_t st = b->stream;
  hipLaunchKernelGGL(k_zero_i32, dim3(64), dim3(256), 0, st,
                     reinterpret_cast<int32_t*>(b->stats.p), (size_t)8 * L * P);
  hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, st, b->iter.p, 0u);
  for (int q = 0; q < P; ++q)
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, st, b->clk.p + 1 + q, ~0ull);
  hipLaunchKernelGGL(k_clock, dim3(1), dim3(1), 0, st, b->clk.p);
  int rc;
  for (int q = 0; q < P; ++q) {
    hea_problem p = b->probs[q];
    dim3 g((p->hc.N + 255) / 256, L);
    hipLaunchKernelGGL(k_permute, g, dim3(256), 0, st, p->d_const.p, b->A.p + b->lat_off[q],
                       b->seed + (uint64_t)q, b->iter.p, (uint32_t)HEA_PHASE_INIT, 0);
    if ((rc = launch_recount(p, b->A.p + b->lat_off[q], b->counts.p + b->cnt_off[q], L, st)))
      return rc;
    if ((rc = launch_fitness(p, b->counts.p + b->cnt_off[q], b->fA.p + (size_t)q * L, L, st)))
      return rc;
  }
  HEA_HIP(hipGetLastError());
  return HEA_OK;
}

int hea_batch_step(hea_batch b, int32_t n_iter) {
  if (!b || n_iter < 0) return fail(HEA_EINVAL, "bad arguments");
  HEA_HIP(hipSetDevice(b->probs[0]->device));
  if (!b->exec) {
    HEA_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = batch_enqueue_iteration(b);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return fail(HEA_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
    b->graph = g;
    HEA_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  }
  for (int i = 0; i < n_iter; ++i) HEA_HIP(hipGraphLaunch(b->exec, b->stream));
  return HEA_OK;
}

int hea_batch_step_timed(hea_batch b, int32_t n_iter, float* out_ms, hea_stats* out_stats) {
  if (!b || n_iter < 0 || !out_ms || !out_stats) return fail(HEA_EINVAL, "bad arguments");
  HEA_HIP(hipSetDevice(b->probs[0]->device));
  const size_t sn = (size_t)4 * b->L * b->probs.size();
  IterHooks hk;
  for (auto& e : hk.ev) HEA_HIP(hipEventCreate(&e));
  DevBuf<int64_t> snap;
  std::vector<int64_t> h(3 * sn);
  int rc = HEA_OK;
  hipError_t e = snap.alloc(3 * sn);
  if (e != hipSuccess) rc = fail(HEA_ENOMEM, hipGetErrorString(e));
  hk.snap = snap.p;
  for (int i = 0; i < n_iter && rc == HEA_OK; ++i) {
    if ((rc = batch_enqueue_iteration(b, &hk))) break;
    e = hipStreamSynchronize(b->stream);
    if (e == hipSuccess) e = hipEventElapsedTime(&out_ms[2 * i], hk.ev[0], hk.ev[1]);
    if (e == hipSuccess) e = hipEventElapsedTime(&out_ms[2 * i + 1], hk.ev[2], hk.ev[3]);
    if (e == hipSuccess)
      e = hipMemcpy(h.data(), snap.p, sizeof(int64_t) * h.size(), hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
      rc = fail(HEA_EHIP, hipGetErrorString(e));
      break;
    }
    for (int ph = 0; ph < 2; ++ph) {
      const std::vector<int64_t> d0(h.begin() + ph * sn, h.begin() + (ph + 1) * sn);
      const std::vector<int64_t> d1(h.begin() + (ph + 1) * sn, h.begin() + (ph + 2) * sn);
      hea_stats a, c;
      sum_stats(d0, (int)(sn / 4), &a);
      sum_stats(d1, (int)(sn / 4), &c);
      out_stats[2 * i + ph] = {c.evals - a.evals, c.rounds - a.rounds, c.accepts - a.accepts,
                               c.draws - a.draws, -1.0};
    }
  }
  for (auto& ev : hk.ev) (void)hipEventDestroy(ev);
  return rc;
}

int hea_batch_sync(hea_batch b) {
  if (!b) return fail(HEA_EINVAL, "null batch");
  HEA_HIP(hipSetDevice(b->probs[0]->device));
  HEA_HIP(hipStreamSynchronize(b->stream));
  return HEA_OK;
}

int hea_batch_read(hea_batch b, int32_t q, uint8_t* out_lat, double* out_fit, hea_stats* stats) {
  if (!b || q < 0 || q >= (int)b->probs.size()) return fail(HEA_EINVAL, "bad arguments");
  hea_problem p = b->probs[q];
  HEA_HIP(hipSetDevice(p->device));
  HEA_HIP(hipStreamSynchronize(b->stream));
  const int L = b->L;
  if (out_lat)
    HEA_HIP(hipMemcpy(out_lat, b->A.p + b->lat_off[q], (size_t)L * p->hc.N,
                      hipMemcpyDeviceToHost));
  if (out_fit)
    HEA_HIP(hipMemcpy(out_fit, b->fA.p + (size_t)q * L, sizeof(double) * L,
                      hipMemcpyDeviceToHost));
  if (stats) {
    std::vector<int64_t> st((size_t)4 * L);
    HEA_HIP(hipMemcpy(st.data(), b->stats.p + (size_t)q * 4 * L, sizeof(int64_t) * st.size(),
                      hipMemcpyDeviceToHost));
    sum_stats(st, L, stats);
    unsigned long long c0 = 0, c1 = 0;
    HEA_HIP(hipMemcpy(&c0, b->clk.p, sizeof(c0), hipMemcpyDeviceToHost));
    HEA_HIP(hipMemcpy(&c1, b->clk.p + 1 + q, sizeof(c1), hipMemcpyDeviceToHost));
    int khz = 0;
    HEA_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, p->device));
    if (c1 != ~0ull && khz > 0) stats->first_hit_s = (double)(c1 - c0) / (1000.0 * khz);
  }
  return HEA_OK;
}

int hea_batch_destroy(hea_batch b) {
  if (!b) return HEA_OK;
  (void)hipSetDevice(b->probs[0]->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return HEA_OK;
}

}  // extern "C"

// ------------------------------------------------------------- relaxation --
hipError_t hea_relax_forward_impl(const HeaConst* dc, const HeaConst& hc, const int32_t* nbrT,
                                  int dtype, const void* th, void* p, void* Q, double* part,
                                  double* n, hipStream_t st);
hipError_t hea_relax_backward_impl(const HeaConst* dc, const HeaConst& hc, int dtype,
                                   const void* p, const void* Q, const double* dn, void* grad,
                                   hipStream_t st);
hipError_t hea_site_gain_impl(const HeaConst* dc, const HeaConst& hc, const int32_t* nbrT,
                              const uint8_t* lattices, int L, const double* dn, float* u,
                              hipStream_t st);
size_t hea_relax_part_size(int n_sites, int K, int S);
hipError_t hea_relax_transpose_nbr(const int32_t* nbr, int N, int stride, int32_t* nbrT,
                                   hipStream_t st);

// First relaxation call on a problem: work buffers and the slot-major copy
// of the neighbour table (waits for `st` once).
static int relax_prepare(hea_problem p, hipStream_t st) {
  if (p->d_relax_nbrT.p) return HEA_OK;
  const HeaConst& c = p->hc;
  HEA_HIP(p->d_relax_part.alloc(hea_relax_part_size(c.N, c.K, c.S)));
  DevBuf<int32_t> t;
  HEA_HIP(t.alloc((size_t)c.N * c.stride));
  HEA_HIP(hea_relax_transpose_nbr(p->d_nbr.p, c.N, c.stride, t.p, st));
  HEA_HIP(hipStreamSynchronize(st));
  std::swap(p->d_relax_nbrT.p, t.p);
  std::swap(p->d_relax_nbrT.n, t.n);
  return HEA_OK;
}

extern "C" {

// Padded p and q rows are read and written as 16-byte vectors.
static bool aligned16(const void* a) { return (reinterpret_cast<uintptr_t>(a) & 15) == 0; }

int hea_relax_forward(hea_problem p, int32_t dtype, const void* d_logits, void* d_p, void* d_q,
                      void* d_n, void* stream) {
  if (!p || !d_logits || !d_p || !d_q || !d_n || (dtype != 0 && dtype != 1))
    return fail(HEA_EINVAL, "bad arguments");
  if (!aligned16(d_p) || !aligned16(d_q)) return fail(HEA_EINVAL, "p and q must be 16-byte aligned");
  HEA_HIP(hipSetDevice(p->device));
  const HeaConst& c = p->hc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (int rc = relax_prepare(p, st)) return rc;
  HEA_HIP(hea_relax_forward_impl(p->d_const.p, c, p->d_relax_nbrT.p, dtype, d_logits, d_p, d_q,
                                 p->d_relax_part.p, static_cast<double*>(d_n), st));
  return HEA_OK;
}

int hea_relax_backward(hea_problem p, int32_t dtype, const void* d_p, const void* d_q,
                       const void* d_dn, void* d_grad, void* stream) {
  if (!p || !d_p || !d_q || !d_dn || !d_grad || (dtype != 0 && dtype != 1))
    return fail(HEA_EINVAL, "bad arguments");
  if (!aligned16(d_p) || !aligned16(d_q)) return fail(HEA_EINVAL, "p and q must be 16-byte aligned");
  HEA_HIP(hipSetDevice(p->device));
  HEA_HIP(hea_relax_backward_impl(p->d_const.p, p->hc, dtype, d_p, d_q,
                                  static_cast<const double*>(d_dn), d_grad,
                                  static_cast<hipStream_t>(stream)));
  return HEA_OK;
}

int hea_site_gain(hea_problem p, int32_t L, const void* d_lattices, const void* d_dn, void* d_u,
                  void* stream) {
  if (!p || L < 1 || !d_lattices || !d_dn || !d_u) return fail(HEA_EINVAL, "bad arguments");
  HEA_HIP(hipSetDevice(p->device));
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (int rc = relax_prepare(p, st)) return rc;
  HEA_HIP(hea_site_gain_impl(p->d_const.p, p->hc, p->d_relax_nbrT.p,
                             static_cast<const uint8_t*>(d_lattices), L,
                             static_cast<const double*>(d_dn), static_cast<float*>(d_u), st));
  return HEA_OK;
}

}  // extern "C"
