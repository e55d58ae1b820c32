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
// Layout: p [N][KP] and Q [S][N][KP] have rows padded with zeros to KP
// columns (KP = 4 for K <= 4, else 8).  The gather gives each site KP / 4
// lanes of four species each, so a neighbour's row is one 4-element vector
// load per lane (16 bytes in float32), and reads the neighbour indices from
// a slot-major copy of the table (nbrT[column][site]), one contiguous run
// per wave; the adjoint reads Q rows and p columns as vectors too.
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

// p[i][c] = softmax(theta[i])[c] for c < K, 0 up to c < KP: KP / 4 lanes
// per site, four columns each, written as one vector.
template <typename T, int KP>
__global__ __launch_bounds__(kSoftmaxThreads) void k_softmax(const T* __restrict__ th,
                                                             T* __restrict__ p, int N, int K) {
  constexpr int GL = KP / 4;
  using Vec = VecT<T, 4>;
  const int t = blockIdx.x * kSoftmaxThreads + threadIdx.x;
  const int i = t / GL, b = t % GL;
  if (i >= N) return;  // a site's lanes leave together
  double x[4], m = -HUGE_VAL;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int c = b * 4 + v;
    x[v] = c < K ? (double)th[(size_t)i * K + c] : -HUGE_VAL;
    m = fmax(m, x[v]);
  }
  if (GL > 1) m = fmax(m, __shfl_xor(m, 1, GL));
  double z = 0.0;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    x[v] = b * 4 + v < K ? exp(x[v] - m) : 0.0;
    z += x[v];
  }
  if (GL > 1) z += __shfl_xor(z, 1, GL);
  Vec out;
#pragma unroll
  for (int v = 0; v < 4; ++v) out.x[v] = (T)(x[v] / z);
  *reinterpret_cast<Vec*>(p + (size_t)i * KP + b * 4) = out;
}

// Q[s][i][c] for one block of SB sites: lane b of site i sums species
// b*V .. b*V+V-1 over i's shell-s neighbours, in list order; then the
// block's share of n[s][a][c] = sum_i p[i][a] Q[s][i][c] from LDS: four
// interleaved site subsets per entry, added in a fixed order.  Rows of pp
// and Q have G*V columns; part is [S K K][nb].
template <typename T, int G, int V>
__global__ __launch_bounds__(kGatherThreads) void k_gather_pairs(
    const HeaConst* __restrict__ pc, const int32_t* __restrict__ nbrT, const T* __restrict__ pp,
    T* __restrict__ Q, double* __restrict__ part) {
  constexpr int SB = kGatherThreads / G, KP = G * V;
  using Vec = VecT<T, V>;
  __shared__ T sp[SB * KP];
  __shared__ T sq[HEA_MAX_S * SB * KP];
  __shared__ double red[4 * HEA_MAX_S * HEA_MAX_K * HEA_MAX_K];
  const HeaConst& P = *pc;
  const int N = P.N, K = P.K, S = P.S;
  const int nb = gridDim.x, blk = xcd_block(blockIdx.x, nb);
  const int tl = threadIdx.x, il = tl / G, b = tl % G;
  const int i = blk * SB + il;
  T* myp = sp + il * KP + b * V;
  if (i < N) {
    const Vec own = *reinterpret_cast<const Vec*>(pp + (size_t)i * KP + b * V);
#pragma unroll
    for (int v = 0; v < V; ++v) myp[v] = own.x[v];
    for (int s = 0; s < S; ++s) {
      const int32_t* col = nbrT + (size_t)P.off[s] * N + i;
      const T* row = pp + b * V;
      const int Z = P.Z[s];
      double acc[V];
#pragma unroll
      for (int v = 0; v < V; ++v) acc[v] = 0.0;
      int k = 0;
      for (; k + 4 <= Z; k += 4) {
        const int j0 = col[(size_t)k * N], j1 = col[(size_t)(k + 1) * N];
        const int j2 = col[(size_t)(k + 2) * N], j3 = col[(size_t)(k + 3) * N];
        const Vec v0 = *reinterpret_cast<const Vec*>(row + (size_t)j0 * KP);
        const Vec v1 = *reinterpret_cast<const Vec*>(row + (size_t)j1 * KP);
        const Vec v2 = *reinterpret_cast<const Vec*>(row + (size_t)j2 * KP);
        const Vec v3 = *reinterpret_cast<const Vec*>(row + (size_t)j3 * KP);
#pragma unroll
        for (int v = 0; v < V; ++v) {
          acc[v] += (double)v0.x[v];
          acc[v] += (double)v1.x[v];
          acc[v] += (double)v2.x[v];
          acc[v] += (double)v3.x[v];
        }
      }
      for (; k < Z; ++k) {
        const Vec vk = *reinterpret_cast<const Vec*>(row + (size_t)col[(size_t)k * N] * KP);
#pragma unroll
        for (int v = 0; v < V; ++v) acc[v] += (double)vk.x[v];
      }
      Vec qv;
#pragma unroll
      for (int v = 0; v < V; ++v) qv.x[v] = (T)acc[v];
      T* myq = sq + s * SB * KP + il * KP + b * V;
#pragma unroll
      for (int v = 0; v < V; ++v) myq[v] = qv.x[v];
      *reinterpret_cast<Vec*>(Q + ((size_t)s * N + i) * KP + b * V) = qv;
    }
  } else {
#pragma unroll
    for (int v = 0; v < V; ++v) myp[v] = (T)0;
    for (int s = 0; s < S; ++s)
#pragma unroll
      for (int v = 0; v < V; ++v) sq[s * SB * KP + il * KP + b * V + v] = (T)0;
  }
  __syncthreads();
  const int E = S * K * K;
  for (int t = tl; t < 4 * E; t += kGatherThreads) {
    const int e = t % E, r = t / E;
    const int s = e / (K * K), a = (e / K) % K, c = e % K;
    const T* q = sq + s * SB * KP;
    double acc = 0.0;
    for (int x = r; x < SB; x += 4) acc += (double)sp[x * KP + a] * (double)q[x * KP + c];
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
// g[k][c] = sum_s sum_b (dn[s][c][b] + dn[s][b][c]) Q[s][k][b]; KP / 4 lanes
// per site, four columns c each.  Every lane loads the site's whole padded
// Q rows and its own four p columns as vectors before the workgroup builds
// the symmetrised table H in LDS, so the loads overlap that step; the two
// halves of <p, g> meet through one shuffle.  grad is [N][K].
template <typename T, int KP>
__global__ __launch_bounds__(256) void k_adjoint(const HeaConst* __restrict__ pc,
                                                 const T* __restrict__ p,
                                                 const T* __restrict__ Q,
                                                 const double* __restrict__ dn,
                                                 T* __restrict__ grad) {
  constexpr int GL = KP / 4;
  using Row = VecT<T, KP>;
  using Vec = VecT<T, 4>;
  const HeaConst& P = *pc;
  const int K = P.K, S = P.S, N = P.N;
  __shared__ double H[HEA_MAX_S * HEA_MAX_K * HEA_MAX_K];
  const int t = blockIdx.x * 256 + threadIdx.x;
  const int k = t / GL, b = t % GL;
  const bool live = k < N;
  Row q[HEA_MAX_S];
  Vec pr;
#pragma unroll
  for (int s = 0; s < HEA_MAX_S; ++s)
    if (live && s < S) q[s] = *reinterpret_cast<const Row*>(Q + ((size_t)s * N + k) * KP);
  if (live) pr = *reinterpret_cast<const Vec*>(p + (size_t)k * KP + b * 4);
  for (int e = threadIdx.x; e < S * K * K; e += blockDim.x) {
    const int s = e / (K * K), c = (e / K) % K, bb = e % K;
    H[e] = dn[(s * K + c) * K + bb] + dn[(s * K + bb) * K + c];
  }
  __syncthreads();
  if (!live) return;
  double g[4], dot = 0.0;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int c = b * 4 + v;
    double a = 0.0;
#pragma unroll
    for (int s = 0; s < HEA_MAX_S; ++s)
#pragma unroll
      for (int bb = 0; bb < KP; ++bb)
        if (c < K && s < S && bb < K) a += H[(s * K + c) * K + bb] * (double)q[s].x[bb];
    g[v] = a;
    if (c < K) dot += (double)pr.x[v] * a;
  }
  if (GL > 1) dot += __shfl_xor(dot, 1, GL);
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int c = b * 4 + v;
    if (c < K) grad[(size_t)k * K + c] = (T)((double)pr.x[v] * (g[v] - dot));
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

// Columns of the padded occupancy rows.
inline int relax_cols(int K) { return K <= 4 ? 4 : 8; }

// Workgroups of the gather (sites per workgroup: kGatherThreads / (KP / 4)).
inline int relax_gather_blocks(int N, int K) {
  const int sb = kGatherThreads / (relax_cols(K) / kGatherVec);
  return (N + sb - 1) / sb;
}

template <typename T, int KP>
hipError_t relax_forward_g(const HeaConst* dc, const HeaConst& hc, const int32_t* nbrT,
                           const void* th, void* p, void* Q, double* part, double* n,
                           hipStream_t st) {
  constexpr int SBS = kSoftmaxThreads / (KP / 4);
  hipLaunchKernelGGL((k_softmax<T, KP>), dim3((hc.N + SBS - 1) / SBS), dim3(kSoftmaxThreads), 0,
                     st, static_cast<const T*>(th), static_cast<T*>(p), hc.N, hc.K);
  const int nb = relax_gather_blocks(hc.N, hc.K);
  hipLaunchKernelGGL((k_gather_pairs<T, KP / kGatherVec, kGatherVec>), dim3(nb),
                     dim3(kGatherThreads), 0, st, dc, nbrT, static_cast<const T*>(p),
                     static_cast<T*>(Q), part);
  hipLaunchKernelGGL(k_pair_reduce, dim3(hc.S * hc.K * hc.K), dim3(256), 0, st, part, nb, n);
  return hipGetLastError();
}

template <typename T, int KP>
hipError_t relax_backward_t(const HeaConst* dc, const HeaConst& hc, const void* p, const void* Q,
                            const double* dn, void* grad, hipStream_t st) {
  const int lanes = hc.N * (KP / 4);
  hipLaunchKernelGGL((k_adjoint<T, KP>), dim3((lanes + 255) / 256), dim3(256), 0, st, dc,
                     static_cast<const T*>(p), static_cast<const T*>(Q), dn,
                     static_cast<T*>(grad));
  return hipGetLastError();
}

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
  const bool g4 = relax_cols(hc.K) == 4;
  if (dtype == 1)
    return g4 ? relax_forward_g<double, 4>(dc, hc, nbrT, th, p, Q, part, n, st)
              : relax_forward_g<double, 8>(dc, hc, nbrT, th, p, Q, part, n, st);
  return g4 ? relax_forward_g<float, 4>(dc, hc, nbrT, th, p, Q, part, n, st)
            : relax_forward_g<float, 8>(dc, hc, nbrT, th, p, Q, part, n, st);
}

hipError_t hea_relax_backward_impl(const HeaConst* dc, const HeaConst& hc, int dtype,
                                   const void* p, const void* Q, const double* dn, void* grad,
                                   hipStream_t st) {
  const bool g4 = relax_cols(hc.K) == 4;
  if (dtype == 1)
    return g4 ? relax_backward_t<double, 4>(dc, hc, p, Q, dn, grad, st)
              : relax_backward_t<double, 8>(dc, hc, p, Q, dn, grad, st);
  return g4 ? relax_backward_t<float, 4>(dc, hc, p, Q, dn, grad, st)
            : relax_backward_t<float, 8>(dc, hc, p, Q, dn, grad, st);
}

hipError_t hea_site_gain_impl(const HeaConst* dc, const HeaConst& hc, const int32_t* nbrT,
                              const uint8_t* lattices, int L, const double* dn, float* u,
                              hipStream_t st) {
  hipLaunchKernelGGL(k_site_gain, dim3((hc.N + 255) / 256, L), dim3(256), 0, st, dc, nbrT,
                     lattices, dn, u);
  return hipGetLastError();
}
