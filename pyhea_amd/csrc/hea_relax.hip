// Differentiable relaxation of the lattice search and its adjoint.
//
// No counterpart in the reference (pyhea/cpp/accelerate.cpp searches only
// discrete lattices).  Each site i carries logits theta[i][K]; its occupancy
// is p[i] = softmax(theta[i]).  The soft pair counts
//     n[s][a][b] = sum_i sum_{j in N_s(i)} p[i][a] p[j][b]
// reduce to the integer pair counts of calculate_sro_coefficient
// (accelerate.cpp:370-399, times Z_s N) when p is one-hot, so
// gamma = n / (Z_s N) and alpha = 1 - gamma / (c_a c_b) relax the
// Warren-Cowley fitness smoothly.
//
// Forward: softmax, neighbour gather Q_s[i][b] = sum_{j in N_s(i)} p[j][b],
// then n_s = P^T Q_s, reduced deterministically in FP64 (a fixed-order sum
// per block of sites, then the block partials in a fixed-shape tree).
//
// Layout of p and Q.  K <= 4: rows of 4 columns, p [N][4], Q [S][N][4]
// (zeros beyond K).  K > 4: each row is split into part A, columns 0-3, and
// part B, columns 4 .. 4 + W2 - 1 with W2 = 1, 2, 4 for K = 5, 6, 7-8 (zeros
// beyond K); p holds part A of every site [N][4], then part B [N][W2]; Q
// holds the shells' A parts [S][N][4], then their B parts [S][N][W2].  A
// site's p row is 4 + W2 elements (hea_relax_cols), so K = 5 moves 5 values
// per row, not 8.  The gather reads a neighbour's part as one vector load
// (16 bytes for part A in float32), with the lanes of a wave all on the
// same part, and reads the neighbour indices from a slot-major copy of the
// table (nbrT[column][site]), one contiguous run per wave; the adjoint reads
// Q and p rows as vectors too.
// Backward (neighbour-gather adjoint): the neighbour relation is symmetric,
// so dn[s][a][b]/dp[k][c] = [a=c] Q_s[k][b] + [b=c] Q_s[k][a] and
//     g[k][c] = sum_s sum_b (dn[s][c][b] + dn[s][b][c]) Q_s[k][b],
// followed by the softmax adjoint p (g - <p, g>).
//
// Guidance at a discrete lattice: with p one-hot, Q_s[k][b] is the integer
// count of species b among k's shell-s neighbours and g[k][c] is the
// first-order change of the loss per unit of occupancy c at site k.
// k_site_gain writes u[k][b] = g[k][b] - g[k][a_k], the predicted loss change
// of turning site k from its species a_k into b.

#include "hea_internal.h"

namespace {

constexpr int kSoftmaxThreads = 512;  // softmax lanes per workgroup
constexpr int kGatherThreads = 512;   // gather lanes per workgroup
constexpr int kGatherVec = 4;         // species per gather lane

// Workgroups are dispatched round-robin over the eight XCDs; this bijection
// of [0, nb) gives each XCD one contiguous range of logical blocks, so the
// neighbour rows its sites read stay in its own L2.
__device__ __forceinline__ int xcd_block(int w, int nb) {
  const int q = nb >> 3, r = nb & 7, x = w & 7, i = w >> 3;
  return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// nbrT[c][i] = nbr[i][c]: the slot-major copy of the neighbour table.
__global__ void k_transpose_nbr(const int32_t* __restrict__ nbr, int N, int stride,
                                int32_t* __restrict__ nbrT) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)N * stride) return;
  const int c = (int)(t / N), i = (int)(t % N);
  nbrT[t] = nbr[(size_t)i * stride + c];
}

// V elements loaded or stored together.  Alignment at most 16 bytes: the
// entry points check 16-byte alignment of p and q, every row starts at a
// multiple of 16 bytes, and wider vectors are moved as 16-byte pieces.
template <typename T, int V>
struct alignas(V * sizeof(T) < 16 ? V * sizeof(T) : 16) VecT {
  T x[V];
};

// Width of part B of the rows (file header): 0 for K <= 4, else the columns
// past the first four rounded up to 1, 2 or 4.
constexpr int relax_w2(int K) { return K <= 4 ? 0 : K == 5 ? 1 : K == 6 ? 2 : 4; }

// p = softmax(theta) in the row layout of the file header: one lane per
// site, part A as one 4-column vector, part B (W2 columns) as another;
// columns from K on hold 0.
template <typename T, int W2>
__global__ __launch_bounds__(kSoftmaxThreads) void k_softmax(const T* __restrict__ th,
                                                             T* __restrict__ p, int N, int K) {
  constexpr int C = 4 + W2;
  const int i = blockIdx.x * kSoftmaxThreads + threadIdx.x;
  if (i >= N) return;
  double x[C], m = -HUGE_VAL;
#pragma unroll
  for (int c = 0; c < C; ++c) {
    x[c] = c < K ? (double)th[(size_t)i * K + c] : -HUGE_VAL;
    m = fmax(m, x[c]);
  }
  double z = 0.0;
#pragma unroll
  for (int c = 0; c < C; ++c) {
    x[c] = c < K ? exp(x[c] - m) : 0.0;
    z += x[c];
  }
  VecT<T, 4> a;
#pragma unroll
  for (int v = 0; v < 4; ++v) a.x[v] = (T)(x[v] / z);
  *reinterpret_cast<VecT<T, 4>*>(p + (size_t)i * 4) = a;
  if constexpr (W2 > 0) {
    VecT<T, W2> bb;
#pragma unroll
    for (int v = 0; v < W2; ++v) bb.x[v] = (T)(x[4 + v] / z);
    *reinterpret_cast<VecT<T, W2>*>(p + (size_t)4 * N + (size_t)i * W2) = bb;
  }
}

// One gather lane: site i's own occupancies of one row part (V columns,
// rows of V elements from `rows`) into LDS, and for every shell the sum of
// its neighbours' rows in list order into LDS and Q (part rows of V from
// `qrows`, shell s at qrows + s N V).
template <typename T, int V, int KPL>
__device__ __forceinline__ void gather_part(const HeaConst& P, const int32_t* __restrict__ nbrT,
                                            const T* __restrict__ rows, T* __restrict__ qrows,
                                            int i, int SB, int il, int col0, T* sp, T* sq) {
  using Vec = VecT<T, V>;
  const int N = P.N, S = P.S;
  T* myp = sp + il * KPL + col0;
  const Vec own = *reinterpret_cast<const Vec*>(rows + (size_t)i * V);
#pragma unroll
  for (int v = 0; v < V; ++v) myp[v] = own.x[v];
  for (int s = 0; s < S; ++s) {
    const int32_t* col = nbrT + (size_t)P.off[s] * N + i;
    const int Z = P.Z[s];
    double acc[V];
#pragma unroll
    for (int v = 0; v < V; ++v) acc[v] = 0.0;
    int k = 0;
    for (; k + 4 <= Z; k += 4) {
      const int j0 = col[(size_t)k * N], j1 = col[(size_t)(k + 1) * N];
      const int j2 = col[(size_t)(k + 2) * N], j3 = col[(size_t)(k + 3) * N];
      const Vec v0 = *reinterpret_cast<const Vec*>(rows + (size_t)j0 * V);
      const Vec v1 = *reinterpret_cast<const Vec*>(rows + (size_t)j1 * V);
      const Vec v2 = *reinterpret_cast<const Vec*>(rows + (size_t)j2 * V);
      const Vec v3 = *reinterpret_cast<const Vec*>(rows + (size_t)j3 * V);
#pragma unroll
      for (int v = 0; v < V; ++v) {
        acc[v] += (double)v0.x[v];
        acc[v] += (double)v1.x[v];
        acc[v] += (double)v2.x[v];
        acc[v] += (double)v3.x[v];
      }
    }
    for (; k < Z; ++k) {
      const Vec vk = *reinterpret_cast<const Vec*>(rows + (size_t)col[(size_t)k * N] * V);
#pragma unroll
      for (int v = 0; v < V; ++v) acc[v] += (double)vk.x[v];
    }
    Vec qv;
#pragma unroll
    for (int v = 0; v < V; ++v) qv.x[v] = (T)acc[v];
    T* myq = sq + s * SB * KPL + il * KPL + col0;
#pragma unroll
    for (int v = 0; v < V; ++v) myq[v] = qv.x[v];
    *reinterpret_cast<Vec*>(qrows + ((size_t)s * N + i) * V) = qv;
  }
}

// Q[s][i][c] for one block of SB sites, then the block's share of
// n[s][a][c] = sum_i p[i][a] Q[s][i][c] from LDS: four interleaved site
// subsets per entry, added in a fixed order; part is [S K K][nb].  With a
// part B (W2 > 0) the first half of the workgroup's waves gathers part A
// of SB = 256 sites and the second half part B of the same sites, so every
// wave loads rows of one width.  LDS rows are KPL = 4 or 8 columns wide.
template <typename T, int W2>
__global__ __launch_bounds__(kGatherThreads) void k_gather_pairs(
    const HeaConst* __restrict__ pc, const int32_t* __restrict__ nbrT, const T* __restrict__ pp,
    T* __restrict__ Q, double* __restrict__ part) {
  constexpr int G = W2 > 0 ? 2 : 1, SB = kGatherThreads / G, KPL = 4 * G;
  __shared__ T sp[SB * KPL];
  __shared__ T sq[HEA_MAX_S * SB * KPL];
  __shared__ double red[4 * HEA_MAX_S * HEA_MAX_K * HEA_MAX_K];
  const HeaConst& P = *pc;
  const int N = P.N, K = P.K, S = P.S;
  const int nb = gridDim.x, blk = xcd_block(blockIdx.x, nb);
  const int tl = threadIdx.x, b = tl / SB, il = tl % SB;  // b is uniform per wave
  const int i = blk * SB + il;
  if (i < N) {
    if (b == 0)
      gather_part<T, 4, KPL>(P, nbrT, pp, Q, i, SB, il, 0, sp, sq);
    else if constexpr (W2 > 0)
      gather_part<T, W2, KPL>(P, nbrT, pp + (size_t)4 * N, Q + (size_t)4 * S * N, i, SB, il, 4,
                              sp, sq);
  }
  // zeros for absent sites and the padding columns the pair sums never read
  // beyond K; every LDS column of a present site is written above or here
  const int c0 = b * 4, cw = b ? KPL - 4 : 4, cset = b ? W2 : 4;
  for (int v = (i < N ? cset : 0); v < cw; ++v) {
    sp[il * KPL + c0 + v] = (T)0;
    for (int s = 0; s < S; ++s) sq[s * SB * KPL + il * KPL + c0 + v] = (T)0;
  }
  __syncthreads();
  const int E = S * K * K;
  for (int t = tl; t < 4 * E; t += kGatherThreads) {
    const int e = t % E, r = t / E;
    const int s = e / (K * K), a = (e / K) % K, c = e % K;
    const T* q = sq + s * SB * KPL;
    double acc = 0.0;
    for (int x = r; x < SB; x += 4) acc += (double)sp[x * KPL + a] * (double)q[x * KPL + c];
    red[r * E + e] = acc;
  }
  __syncthreads();
  for (int e = tl; e < E; e += kGatherThreads)
    part[(size_t)e * nb + blk] = ((red[e] + red[E + e]) + red[2 * E + e]) + red[3 * E + e];
}

// n[e] = sum over the nb block partials of entry e: strided lane sums, then
// a fixed-shape tree (one workgroup per entry).
__global__ __launch_bounds__(256) void k_pair_reduce(const double* __restrict__ part, int nb,
                                                     double* __restrict__ n) {
  __shared__ double r[256];
  const double* row = part + (size_t)blockIdx.x * nb;
  double acc = 0.0;
  for (int x = threadIdx.x; x < nb; x += 256) acc += row[x];
  r[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) r[threadIdx.x] += r[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) n[blockIdx.x] = r[0];
}

// grad[k][c] = p[k][c] (g[k][c] - sum_d p[k][d] g[k][d]),
// g[k][c] = sum_s sum_b (dn[s][c][b] + dn[s][b][c]) Q[s][k][b]; one lane
// per site for K <= 4, else two, four columns c each.  Every lane loads the
// site's whole Q rows and p row (both parts, as vectors) before the
// workgroup builds the symmetrised table H in LDS, so the loads overlap that
// step; the two halves of <p, g> meet through one shuffle.  grad is [N][K].
template <typename T, int W2>
__global__ __launch_bounds__(256) void k_adjoint(const HeaConst* __restrict__ pc,
                                                 const T* __restrict__ p,
                                                 const T* __restrict__ Q,
                                                 const double* __restrict__ dn,
                                                 T* __restrict__ grad) {
  constexpr int GL = W2 > 0 ? 2 : 1, C = 4 + W2;
  using VA = VecT<T, 4>;
  using VB = VecT<T, (W2 > 0 ? W2 : 1)>;
  const HeaConst& P = *pc;
  const int K = P.K, S = P.S, N = P.N;
  __shared__ double H[HEA_MAX_S * HEA_MAX_K * HEA_MAX_K];
  const int t = blockIdx.x * 256 + threadIdx.x;
  const int k = t / GL, b = t % GL;
  const bool live = k < N;
  T q[HEA_MAX_S][C], pr[C];
  if (live) {
#pragma unroll
    for (int s = 0; s < HEA_MAX_S; ++s)
      if (s < S) {
        const VA qa = *reinterpret_cast<const VA*>(Q + ((size_t)s * N + k) * 4);
#pragma unroll
        for (int v = 0; v < 4; ++v) q[s][v] = qa.x[v];
        if constexpr (W2 > 0) {
          const VB qb = *reinterpret_cast<const VB*>(Q + (size_t)4 * S * N + ((size_t)s * N + k) * W2);
#pragma unroll
          for (int v = 0; v < W2; ++v) q[s][4 + v] = qb.x[v];
        }
      }
    const VA pa = *reinterpret_cast<const VA*>(p + (size_t)k * 4);
#pragma unroll
    for (int v = 0; v < 4; ++v) pr[v] = pa.x[v];
    if constexpr (W2 > 0) {
      const VB pb = *reinterpret_cast<const VB*>(p + (size_t)4 * N + (size_t)k * W2);
#pragma unroll
      for (int v = 0; v < W2; ++v) pr[4 + v] = pb.x[v];
    }
  }
  for (int e = threadIdx.x; e < S * K * K; e += blockDim.x) {
    const int s = e / (K * K), c = (e / K) % K, bb = e % K;
    H[e] = dn[(s * K + c) * K + bb] + dn[(s * K + bb) * K + c];
  }
  __syncthreads();
  if (!live) return;
  double g[4], pk[4], dot = 0.0;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int c = b * 4 + v;
    double a = 0.0;
#pragma unroll
    for (int s = 0; s < HEA_MAX_S; ++s)
#pragma unroll
      for (int bb = 0; bb < C; ++bb)
        if (c < K && s < S && bb < K) a += H[(s * K + c) * K + bb] * (double)q[s][bb];
    g[v] = a;
    pk[v] = 0.0;
#pragma unroll
    for (int cc = 0; cc < C; ++cc)  // this lane's column c of the p row
      if (cc == c) pk[v] = (double)pr[cc];
    if (c < K) dot += pk[v] * a;
  }
  if (GL > 1) dot += __shfl_xor(dot, 1, GL);
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int c = b * 4 + v;
    if (c < K) grad[(size_t)k * K + c] = (T)(pk[v] * (g[v] - dot));
  }
}

// u[l][k][b] = g[k][b] - g[k][a_k] at the one-hot lattice l (see file header).
__global__ void k_site_gain(const HeaConst* __restrict__ pc, const int32_t* __restrict__ nbrT,
                            const uint8_t* __restrict__ lattices, const double* __restrict__ dn,
                            float* __restrict__ u) {
  const HeaConst& P = *pc;
  const int K = P.K, S = P.S;
  const int l = blockIdx.y;
  __shared__ double H[HEA_MAX_S * HEA_MAX_K * HEA_MAX_K];
  const double* d = dn + (size_t)l * S * K * K;
  for (int e = threadIdx.x; e < S * K * K; e += blockDim.x) {
    const int s = e / (K * K), c = (e / K) % K, b = e % K;
    H[e] = d[(s * K + c) * K + b] + d[(s * K + b) * K + c];
  }
  __syncthreads();
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= P.N) return;
  const uint8_t* sp = lattices + (size_t)l * P.N;
  double g[HEA_MAX_K];
  for (int c = 0; c < K; ++c) g[c] = 0.0;
  for (int s = 0; s < S; ++s)
    for (int q = 0; q < P.Z[s]; ++q) {
      const int b = sp[nbrT[(size_t)(P.off[s] + q) * P.N + k]];
      for (int c = 0; c < K; ++c) g[c] += H[(s * K + c) * K + b];
    }
  const int a = sp[k];
  float* out = u + ((size_t)l * P.N + k) * K;
  for (int b = 0; b < K; ++b) out[b] = (float)(g[b] - g[a]);
}

// Elements of one site's p row (both parts; the file header's layout).
inline int relax_cols(int K) { return 4 + relax_w2(K); }

// Workgroups of the gather (sites per workgroup: 512 for K <= 4, else 256).
inline int relax_gather_blocks(int N, int K) {
  const int sb = kGatherThreads / (relax_w2(K) > 0 ? 2 : 1);
  return (N + sb - 1) / sb;
}

template <typename T, int W2>
hipError_t relax_forward_g(const HeaConst* dc, const HeaConst& hc, const int32_t* nbrT,
                           const void* th, void* p, void* Q, double* part, double* n,
                           hipStream_t st) {
  hipLaunchKernelGGL((k_softmax<T, W2>), dim3((hc.N + kSoftmaxThreads - 1) / kSoftmaxThreads),
                     dim3(kSoftmaxThreads), 0, st, static_cast<const T*>(th), static_cast<T*>(p),
                     hc.N, hc.K);
  const int nb = relax_gather_blocks(hc.N, hc.K);
  hipLaunchKernelGGL((k_gather_pairs<T, W2>), dim3(nb), dim3(kGatherThreads), 0, st, dc, nbrT,
                     static_cast<const T*>(p), static_cast<T*>(Q), part);
  hipLaunchKernelGGL(k_pair_reduce, dim3(hc.S * hc.K * hc.K), dim3(256), 0, st, part, nb, n);
  return hipGetLastError();
}

template <typename T, int W2>
hipError_t relax_backward_t(const HeaConst* dc, const HeaConst& hc, const void* p, const void* Q,
                            const double* dn, void* grad, hipStream_t st) {
  const int lanes = hc.N * (W2 > 0 ? 2 : 1);
  hipLaunchKernelGGL((k_adjoint<T, W2>), dim3((lanes + 255) / 256), dim3(256), 0, st, dc,
                     static_cast<const T*>(p), static_cast<const T*>(Q), dn,
                     static_cast<T*>(grad));
  return hipGetLastError();
}

// Calls F<T, W2>() for the problem's species count and dtype.
template <template <typename, int> class F, typename... Args>
hipError_t relax_dispatch(int dtype, int K, Args... args) {
  const int w2 = relax_w2(K);
  if (dtype == 1) {
    switch (w2) {
      case 0: return F<double, 0>::run(args...);
      case 1: return F<double, 1>::run(args...);
      case 2: return F<double, 2>::run(args...);
      default: return F<double, 4>::run(args...);
    }
  }
  switch (w2) {
    case 0: return F<float, 0>::run(args...);
    case 1: return F<float, 1>::run(args...);
    case 2: return F<float, 2>::run(args...);
    default: return F<float, 4>::run(args...);
  }
}

template <typename T, int W2>
struct RelaxForward {
  static hipError_t run(const HeaConst* dc, const HeaConst* hc, const int32_t* nbrT,
                        const void* th, void* p, void* Q, double* part, double* n,
                        hipStream_t st) {
    return relax_forward_g<T, W2>(dc, *hc, nbrT, th, p, Q, part, n, st);
  }
};

template <typename T, int W2>
struct RelaxBackward {
  static hipError_t run(const HeaConst* dc, const HeaConst* hc, const void* p, const void* Q,
                        const double* dn, void* grad, hipStream_t st) {
    return relax_backward_t<T, W2>(dc, *hc, p, Q, dn, grad, st);
  }
};

}  // namespace

extern "C" int32_t hea_relax_cols(int32_t n_species) {
  return n_species < 1 || n_species > HEA_MAX_K ? -1 : relax_cols(n_species);
}

size_t hea_relax_part_size(int n_sites, int K, int S) {
  return (size_t)relax_gather_blocks(n_sites, K) * S * K * K;
}

hipError_t hea_relax_transpose_nbr(const int32_t* nbr, int N, int stride, int32_t* nbrT,
                                   hipStream_t st) {
  const size_t n = (size_t)N * stride;
  hipLaunchKernelGGL(k_transpose_nbr, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, nbr, N,
                     stride, nbrT);
  return hipGetLastError();
}

hipError_t hea_relax_forward_impl(const HeaConst* dc, const HeaConst& hc, const int32_t* nbrT,
                                  int dtype, const void* th, void* p, void* Q, double* part,
                                  double* n, hipStream_t st) {
  return relax_dispatch<RelaxForward>(dtype, hc.K, dc, &hc, nbrT, th, p, Q, part, n, st);
}

hipError_t hea_relax_backward_impl(const HeaConst* dc, const HeaConst& hc, int dtype,
                                   const void* p, const void* Q, const double* dn, void* grad,
                                   hipStream_t st) {
  return relax_dispatch<RelaxBackward>(dtype, hc.K, dc, &hc, p, Q, dn, grad, st);
}

hipError_t hea_site_gain_impl(const HeaConst* dc, const HeaConst& hc, const int32_t* nbrT,
                              const uint8_t* lattices, int L, const double* dn, float* u,
                              hipStream_t st) {
  hipLaunchKernelGGL(k_site_gain, dim3((hc.N + 255) / 256, L), dim3(256), 0, st, dc, nbrT,
                     lattices, dn, u);
  return hipGetLastError();
}
