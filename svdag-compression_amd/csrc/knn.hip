// k-nearest-neighbour search over 32-d embeddings for the learned lossy-DAG
// proposer: `tables` random-hyperplane LSH orderings; per ordering each point
// compares itself with the points within +-window sorted positions.  Exact,
// deterministic arithmetic (sequential fp32 sums, no contraction) so the CPU
// restatement reproduces the lists bit for bit:
//   key bit b of table t = dot(e, h[t][b]) > 0, h from a SplitMix64 stream
//   (component = ((z >> 40) * 2^-24) * 2 - 1);
//   stable sort by key; distance = sum_d (e_x[d] - e_y[d])^2;
//   result = the k smallest distinct (distance, index) pairs over all tables.
// With signatures given, the candidates are scored by the exact Hamming
// distance of the 512-bit occupancies instead (the embedding only decides
// which nodes meet in a window): the learned proposer's re-ranking.
#include "svd_internal.hpp"
#include <rocprim/device/device_radix_sort.hpp>

namespace svd {

namespace {

constexpr int kDim = 32;
constexpr int kMaxK = 32;
constexpr int kTile = 256;
constexpr int kRow = kDim + 1;             // padded LDS row (no bank conflicts)

__global__ void k_lsh_keys(const float* __restrict__ e, uint64_t n, const float* __restrict__ h, uint32_t bits,
                           uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float x[kDim];
#pragma unroll
  for (int d = 0; d < kDim; ++d) x[d] = e[i * kDim + d];
  uint32_t key = 0;
  for (uint32_t b = 0; b < bits; ++b) {
    const float* hp = h + (uint64_t)b * kDim;
    float s = 0.f;
#pragma unroll
    for (int d = 0; d < kDim; ++d) { const float q = x[d] * hp[d]; s = s + q; }
    if (s > 0.f) key |= 1u << b;
  }
  keys[i] = key;
  vals[i] = (uint32_t)i;
}

// Block = 256 consecutive sorted positions; the window's rows are staged in
// LDS.  Each lane keeps its point's running top-k (registers) across tables
// through best_d / best_i in HBM.  HAM = 0: squared L2 of the embeddings;
// HAM = 1: Hamming distance of the 512-bit signatures.
template <int HAM, int KM>
__global__ void __launch_bounds__(kTile) k_window_topk(const float* __restrict__ e, const uint32_t* __restrict__ sig,
                                                       const uint32_t* __restrict__ order, uint64_t n, uint32_t W,
                                                       uint32_t k, int first, float* __restrict__ best_d,
                                                       uint32_t* __restrict__ best_i) {
  constexpr int kWords = HAM ? 16 : kDim;
  constexpr int kR = kWords + 1;            // padded LDS row (no bank conflicts)
  extern __shared__ float rows[];           // (kTile + 2W) x kR
  uint32_t* urows = reinterpret_cast<uint32_t*>(rows);
  const uint64_t q0 = (uint64_t)blockIdx.x * kTile;
  const int64_t lo = (int64_t)q0 - (int64_t)W;
  const uint32_t nrows = kTile + 2 * W;
  for (uint32_t t = threadIdx.x; t < nrows * kWords; t += kTile) {
    const uint32_t rr = t / kWords, d = t % kWords;
    const int64_t q = lo + (int64_t)rr;
    const bool in = q >= 0 && q < (int64_t)n;
    if (HAM) urows[rr * kR + d] = in ? sig[(uint64_t)order[q] * 16 + d] : 0u;
    else rows[rr * kR + d] = in ? e[(uint64_t)order[q] * kDim + d] : 0.f;
  }
  __syncthreads();
  const uint64_t q = q0 + threadIdx.x;
  if (q >= n) return;
  const uint32_t x = order[q];
  float bd[KM];
  uint32_t bi[KM];
#pragma unroll
  for (int t = 0; t < KM; ++t) {
    bd[t] = first || t >= (int)k ? __builtin_inff() : best_d[(uint64_t)x * k + t];
    bi[t] = first || t >= (int)k ? 0xFFFFFFFFu : best_i[(uint64_t)x * k + t];
  }
  const uint32_t xr = (threadIdx.x + W) * kR;
  // the query row lives in registers
  uint32_t xs[HAM ? 16 : 1];
  float xe[HAM ? 1 : kDim];
  if (HAM) {
#pragma unroll
    for (int d = 0; d < (HAM ? 16 : 1); ++d) xs[d] = urows[xr + d];
  } else {
#pragma unroll
    for (int d = 0; d < (HAM ? 1 : kDim); ++d) xe[d] = rows[xr + d];
  }
  for (int64_t r = (int64_t)q - (int64_t)W; r <= (int64_t)q + (int64_t)W; ++r) {
    if (r < 0 || r >= (int64_t)n || r == (int64_t)q) continue;
    const uint32_t yr = (uint32_t)(r - lo) * kR;
    float s = 0.f;
    if (HAM) {
      uint32_t h = 0;
#pragma unroll
      for (int d = 0; d < (HAM ? 16 : 1); ++d) h += __popc(xs[d] ^ urows[yr + d]);
      s = (float)h;
    } else {
#pragma unroll
      for (int d = 0; d < (HAM ? 1 : kDim); ++d) { const float t = xe[d] - rows[yr + d]; const float sq = t * t; s = s + sq; }
    }
    // (s, y) < worst kept?
    float wd = bd[KM - 1];
    uint32_t wi = bi[KM - 1];
    if (KM != (int)k) {
#pragma unroll
      for (int t = 0; t < KM; ++t)
        if (t == (int)k - 1) { wd = bd[t]; wi = bi[t]; }
    }
    const uint32_t y = order[r];
    if (!(s < wd || (s == wd && y < wi))) continue;
    bool dup = false;
#pragma unroll
    for (int t = 0; t < KM; ++t) dup |= (t < (int)k) && bi[t] == y;
    if (dup) continue;
    // insert keeping (distance, index) ascending
    float cd = s;
    uint32_t ci = y;
#pragma unroll
    for (int t = 0; t < KM; ++t) {
      if (t < (int)k && (cd < bd[t] || (cd == bd[t] && ci < bi[t]))) {
        const float td = bd[t]; const uint32_t ti = bi[t];
        bd[t] = cd; bi[t] = ci;
        cd = td; ci = ti;
      }
    }
  }
#pragma unroll
  for (int t = 0; t < KM; ++t)
    if (t < (int)k) { best_d[(uint64_t)x * k + t] = bd[t]; best_i[(uint64_t)x * k + t] = bi[t]; }
}

// One neighbour-of-neighbour round (NN-descent) on the signature Hamming
// metric: node x keeps the k smallest distinct (distance, index) among its
// own list and the lists of its neighbours.  Reads `in`, writes `out`.
// One wavefront per node: the k + k*k candidate ids are deduplicated through
// a hash set in LDS, each distinct candidate's signature is read once (by the
// lane whose insert won), and k wave-wide minimum reductions over
// (distance << 32 | index) take the k smallest.
template <int KM>
__global__ void __launch_bounds__(256) k_knn_refine_wave(const uint32_t* __restrict__ sig, uint64_t n, uint32_t k,
                                                         const uint32_t* __restrict__ in, uint32_t* __restrict__ out) {
  constexpr int kCand = KM * (KM + 1);
  constexpr int kPer = (kCand + 63) / 64;                 // candidates per lane
  constexpr int kSlots = KM <= 16 ? 1024 : 4096;          // >= 2 * kCand, power of two
  constexpr int kShift = KM <= 16 ? 22 : 20;              // 32 - log2(kSlots): the product's high bits
  __shared__ uint32_t set[4][kSlots];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint64_t x = (uint64_t)blockIdx.x * 4 + wave;
  const bool live = x < n;
  uint32_t* S = set[wave];
  for (int j = lane; j < kSlots; j += 64) S[j] = 0xFFFFFFFFu;
  __syncthreads();
  const uint32_t nc = k + k * k;
  uint64_t key[kPer];
#pragma unroll
  for (int p = 0; p < kPer; ++p) {
    key[p] = ~0ull;
    const uint32_t c = (uint32_t)(lane + 64 * p);
    if (!live || c >= nc) continue;
    uint32_t z;
    if (c < k) {
      z = in[x * k + c];
    } else {
      const uint32_t y = in[x * k + (c - k) / k];
      z = (y == 0xFFFFFFFFu || (uint64_t)y >= n) ? 0xFFFFFFFFu : in[(uint64_t)y * k + (c - k) % k];
    }
    if (z == 0xFFFFFFFFu || (uint64_t)z == x || (uint64_t)z >= n) continue;
    uint32_t h = (z * 2654435761u) >> kShift;
    bool mine = false;
    while (true) {
      const uint32_t old = atomicCAS(&S[h], 0xFFFFFFFFu, z);
      if (old == 0xFFFFFFFFu) { mine = true; break; }
      if (old == z) break;
      h = (h + 1) & (kSlots - 1);
    }
    if (!mine) continue;
    const uint4* px = reinterpret_cast<const uint4*>(sig + x * 16);
    const uint4* pz = reinterpret_cast<const uint4*>(sig + (uint64_t)z * 16);
    uint32_t d = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint4 a = px[q], b = pz[q];
      d += __popc(a.x ^ b.x) + __popc(a.y ^ b.y) + __popc(a.z ^ b.z) + __popc(a.w ^ b.w);
    }
    key[p] = ((uint64_t)d << 32) | z;
  }
  for (uint32_t t = 0; t < k; ++t) {
    uint64_t m = key[0];
#pragma unroll
    for (int p = 1; p < kPer; ++p) m = key[p] < m ? key[p] : m;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const uint32_t lo = __shfl_xor((uint32_t)m, off, 64), hi = __shfl_xor((uint32_t)(m >> 32), off, 64);
      const uint64_t o = ((uint64_t)hi << 32) | lo;
      m = o < m ? o : m;
    }
    if (live && lane == 0) out[x * k + t] = m == ~0ull ? 0xFFFFFFFFu : (uint32_t)m;
#pragma unroll
    for (int p = 0; p < kPer; ++p)
      if (key[p] == m) key[p] = ~0ull;   // candidates are distinct: one lane holds m
  }
}

}  // namespace

void knn_refine(svd_ctx* c, const uint32_t* sig, uint64_t n, uint32_t k, uint32_t rounds, uint32_t* knn) {
  if (!rounds || !n) return;
  if (k < 1 || k > (uint32_t)kMaxK) throw Error{SVD_ERR_ARG, "knn: k must be 1..32"};
  uint32_t* tmp = dalloc<uint32_t>(c, n * k);
  for (uint32_t r = 0; r < rounds; ++r) {
    if (k <= 16) k_knn_refine_wave<16><<<div_up(n, 4), 256, 0, c->stream>>>(sig, n, k, knn, tmp);
    else k_knn_refine_wave<kMaxK><<<div_up(n, 4), 256, 0, c->stream>>>(sig, n, k, knn, tmp);
    SVD_HIP(hipMemcpyAsync(knn, tmp, n * k * 4, hipMemcpyDeviceToDevice, c->stream));
  }
  check_launch("k_knn_refine");
  dfree(c, tmp);
}

void knn_lsh(svd_ctx* c, const float* emb, uint64_t n, uint32_t k, uint32_t tables, uint32_t bits, uint32_t window,
             uint64_t seed, uint32_t* out, const uint32_t* sig) {
  if (k < 1 || k > (uint32_t)kMaxK) throw Error{SVD_ERR_ARG, "knn: k must be 1..32"};
  if (bits < 1 || bits > 24 || tables < 1 || window < 1 || window > 96) throw Error{SVD_ERR_ARG, "knn: bits 1..24, tables >= 1, window 1..96"};
  if (n >= 0xFFFFFFFFull) throw Error{SVD_ERR_CAPACITY, "knn: too many points"};
  if (n == 0) return;
  // hyperplanes of every table
  std::vector<float> h((size_t)tables * bits * kDim);
  uint64_t s = seed;
  for (auto& v : h) {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    v = (float)(z >> 40) * (1.0f / 16777216.0f) * 2.0f - 1.0f;
  }
  float* dh = dalloc<float>(c, h.size());
  SVD_HIP(hipMemcpyAsync(dh, h.data(), h.size() * 4, hipMemcpyHostToDevice, c->stream));
  uint32_t* keys = dalloc<uint32_t>(c, n);
  uint32_t* keys2 = dalloc<uint32_t>(c, n);
  uint32_t* vals = dalloc<uint32_t>(c, n);
  uint32_t* order = dalloc<uint32_t>(c, n);
  float* bd = dalloc<float>(c, n * k);
  size_t bytes = 0;
  SVD_HIP(rocprim::radix_sort_pairs(nullptr, bytes, keys, keys2, vals, order, (size_t)n, 0, bits, c->stream));
  void* tmp = c->arena->get(bytes + 16);
  const size_t lds = (size_t)(kTile + 2 * window) * (sig ? 17 : kRow) * sizeof(float);
  for (uint32_t t = 0; t < tables; ++t) {
    k_lsh_keys<<<div_up(n, 256), 256, 0, c->stream>>>(emb, n, dh + (size_t)t * bits * kDim, bits, keys, vals);
    SVD_HIP(rocprim::radix_sort_pairs(tmp, bytes, keys, keys2, vals, order, (size_t)n, 0, bits, c->stream));
    const dim3 g(div_up(n, kTile));
    if (sig && k == 16) k_window_topk<1, 16><<<g, kTile, lds, c->stream>>>(emb, sig, order, n, window, k, t == 0, bd, out);
    else if (sig) k_window_topk<1, kMaxK><<<g, kTile, lds, c->stream>>>(emb, sig, order, n, window, k, t == 0, bd, out);
    else if (k == 16) k_window_topk<0, 16><<<g, kTile, lds, c->stream>>>(emb, sig, order, n, window, k, t == 0, bd, out);
    else k_window_topk<0, kMaxK><<<g, kTile, lds, c->stream>>>(emb, sig, order, n, window, k, t == 0, bd, out);
    check_launch("knn table");
  }
  c->arena->put(tmp);
  dfree(c, dh); dfree(c, keys); dfree(c, keys2); dfree(c, vals); dfree(c, order); dfree(c, bd);
}

}  // namespace svd
