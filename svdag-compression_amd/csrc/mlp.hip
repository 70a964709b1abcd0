// Learned candidate encoder for the lossy DAG: MLP 512 -> 256 -> 128 -> 32.
// bf16 weights and activations, fp32 accumulation on the CDNA matrix cores
// (v_mfma_f32_32x32x8_bf16), hand-written forward and backward passes, Adam.
// The reference has no learned part (its lossy path clusters an exhaustive
// similarity graph with MCL, geom_octree_extension.cpp:797-1179).
//
// Definitions (also stated by the CPU oracle the tests compare with):
//   z1 = x W1 + b1, h1 = bf16(relu(z1)); z2 = h1 W2 + b2, h2 = bf16(relu(z2));
//   e = h2 W3 + b3 (fp32).  x = the node's 512-bit occupancy, expanded to
//   0/1 inside the first GEMM's operand load (no expanded copy in HBM).
//   Pair p = batch rows (2p, 2p+1); d = |e_a - e_b|^2; y = min(H, 32);
//   loss = sum (d - y)^2 / P.  Each layer's delta is rounded to bf16 once and
//   feeds both its weight gradient and its bias gradient.
#include <algorithm>
#include "svd_internal.hpp"
#include <rocprim/device/device_radix_sort.hpp>
#include <cmath>

namespace svd {

namespace {

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

constexpr int D0 = 512, D1 = 256, D2 = 128, D3 = 32;
constexpr uint64_t OFF_W1 = 0, OFF_B1 = OFF_W1 + D0 * D1, OFF_W2 = OFF_B1 + D1, OFF_B2 = OFF_W2 + D1 * D2,
                   OFF_W3 = OFF_B2 + D2, OFF_B3 = OFF_W3 + D2 * D3, N_PARAMS = OFF_B3 + D3;
constexpr uint64_t WB_W1 = 0, WB_W2 = WB_W1 + D0 * D1, WB_W3 = WB_W2 + D1 * D2, N_WB = WB_W3 + D2 * D3;
constexpr float kHCap = 32.f;
static_assert(N_PARAMS == SVD_MLP_PARAMS, "parameter layout");

__device__ __forceinline__ uint16_t f2bf(float f) {   // round to nearest even (finite values)
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

enum AMode { A_BITS = 0, A_BITS_T = 1, A_BF16 = 2 };
enum Epi { E_BIAS_RELU_BF16 = 0, E_BIAS_F32 = 1, E_MASK_BF16 = 2, E_F32 = 3 };

struct GemmArgs {
  uint64_t mValid;                 // rows of C that exist (grid may pad to 32)
  int N, K;
  const uint32_t* abits;           // A_BITS: A[m][k] = bit k of signature m; A_BITS_T: bit m of signature k
  const uint16_t* a; int64_t a_sm, a_sk;   // A_BF16: A[m][k] = a[m*a_sm + k*a_sk]
  const uint16_t* b; int64_t b_sk, b_sn;   // B[k][n] = b[k*b_sk + n*b_sn]
  const float* bias;
  const uint16_t* mask; int64_t mask_sm;   // E_MASK_BF16: keep C[m][n] where mask[m*mask_sm + n] > 0
  uint16_t* outb; float* outf; int64_t o_sm;
  int kPer; int64_t o_sz;          // split-K: blockIdx.z covers k in [z*kPer, (z+1)*kPer), writes outf + z*o_sz
};

// One wavefront computes one 32x32 tile of C = A(MxK) B(KxN) with
// v_mfma_f32_32x32x8_bf16: lane l holds A[m0 + l%32][k0 + 4(l/32) + i] and
// B[k0 + 4(l/32) + i][n0 + l%32], i = 0..3; accumulator element v of lane l is
// C[m0 + 8(v/4) + 4(l/32) + v%4][n0 + l%32].
template <int AM, int EP>
__global__ void __launch_bounds__(64) k_gemm(GemmArgs g) {
  const int lane = threadIdx.x;
  const uint64_t m0 = (uint64_t)blockIdx.x * 32;
  const int n0 = blockIdx.y * 32;
  const int r = lane & 31, kq = (lane >> 5) * 4;
  const uint64_t am = m0 + (uint64_t)r;
  const int bn = n0 + r;
  f32x16_t acc;
#pragma unroll
  for (int v = 0; v < 16; ++v) acc[v] = 0.f;
  const int kBeg = (int)blockIdx.z * g.kPer;
  const int kEnd = min(g.K, kBeg + g.kPer);
  for (int k0 = kBeg; k0 < kEnd; k0 += 8) {
    bf16x4_t a, b;
    const int kb = k0 + kq;
    if (AM == A_BITS) {
      const uint32_t w = am < g.mValid ? g.abits[am * 16 + (kb >> 5)] >> (kb & 31) : 0u;
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = ((w >> i) & 1u) ? (short)0x3F80 : (short)0;
    } else if (AM == A_BITS_T) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
        a[i] = ((g.abits[(uint64_t)(kb + i) * 16 + (am >> 5)] >> (am & 31)) & 1u) ? (short)0x3F80 : (short)0;
    } else {
#pragma unroll
      for (int i = 0; i < 4; ++i)
        a[i] = am < g.mValid ? (short)g.a[(int64_t)am * g.a_sm + (int64_t)(kb + i) * g.a_sk] : (short)0;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) b[i] = (short)g.b[(int64_t)(kb + i) * g.b_sk + (int64_t)bn * g.b_sn];
    acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
  }
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    const uint64_t m = m0 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
    const int n = n0 + (lane & 31);
    if (m >= g.mValid) continue;
    const float x = acc[v];
    const int64_t o = (int64_t)m * g.o_sm + n;
    if (EP == E_BIAS_RELU_BF16) {
      const float z = x + g.bias[n];
      g.outb[o] = f2bf(z > 0.f ? z : 0.f);
    } else if (EP == E_BIAS_F32) {
      g.outf[o] = x + g.bias[n];
    } else if (EP == E_MASK_BF16) {
      const uint16_t mk = g.mask[(int64_t)m * g.mask_sm + n];
      g.outb[o] = f2bf(bf2f(mk) > 0.f ? x : 0.f);
    } else {
      g.outf[(int64_t)blockIdx.z * g.o_sz + o] = x;
    }
  }
}

template <int AM, int EP>
void gemm(svd_ctx* c, uint64_t M, int N, int K, const GemmArgs& g0) {
  GemmArgs g = g0;
  g.N = N; g.K = K;
  g.kPer = K;
  g.o_sz = 0;
  dim3 grid(div_up(M, 32), N / 32);
  k_gemm<AM, EP><<<grid, 64, 0, c->stream>>>(g);
}

// out[m][n] = sum over splits s = 0..S-1 (in this order) of part[s][m][n]
__global__ void k_split_sum(const float* __restrict__ part, int S, uint64_t MN, int N, float* __restrict__ out,
                            int64_t o_sm) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= MN) return;
  float t = 0.f;
  for (int s = 0; s < S; ++s) t += part[(uint64_t)s * MN + i];
  out[(int64_t)(i / N) * o_sm + (int64_t)(i % N)] = t;
}

// Weight gradients: M x N is small and K (the rows of the batch) is long, so
// the K range is split over enough workgroups to fill the device; the partial
// tiles are summed in a fixed order (deterministic, no atomics).
constexpr int kSplitTiles = 2048;     // workgroups to aim for
constexpr int kMinKPer = 128;         // at least 16 MFMA steps per workgroup
uint64_t split_part_floats(uint64_t M, int N, int K) {
  const uint64_t tiles = div_up(M, 32) * (uint64_t)(N / 32);
  uint64_t S = std::max<uint64_t>(1, kSplitTiles / tiles);
  S = std::min<uint64_t>(S, std::max(1, K / kMinKPer));
  return S * M * N;
}

template <int AM>
void gemm_wgrad(svd_ctx* c, uint64_t M, int N, int K, const GemmArgs& g0, float* part) {
  GemmArgs g = g0;
  const uint64_t tiles = div_up(M, 32) * (uint64_t)(N / 32);
  int S = (int)std::max<uint64_t>(1, kSplitTiles / tiles);
  S = std::min(S, std::max(1, K / kMinKPer));
  int kPer = (K + S - 1) / S;
  kPer = (kPer + 7) / 8 * 8;
  S = (K + kPer - 1) / kPer;
  float* out = g.outf;
  const int64_t osm = g.o_sm;
  g.N = N; g.K = K;
  g.kPer = kPer;
  g.outf = part;
  g.o_sm = N;
  g.o_sz = (int64_t)M * N;
  dim3 grid(div_up(M, 32), N / 32, S);
  k_gemm<AM, E_F32><<<grid, 64, 0, c->stream>>>(g);
  k_split_sum<<<div_up(M * N, 256), 256, 0, c->stream>>>(part, S, M * N, N, out, osm);
}

// Column sums of a bf16 [R x N] matrix (bias gradients), one block per
// column, fixed reduction tree.
__global__ void __launch_bounds__(256) k_colsum_bf16(const uint16_t* __restrict__ x, uint64_t R, int N,
                                                     float* __restrict__ out) {
  __shared__ float red[256];
  const int n = blockIdx.x;
  float s = 0.f;
  for (uint64_t r = threadIdx.x; r < R; r += 256) s += bf2f(x[r * N + n]);
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[n] = red[0];
}

// Pair loss and its gradient w.r.t. the embeddings (rounded to bf16).
__global__ void k_pair_loss(const float* __restrict__ e, const uint32_t* __restrict__ labels, uint64_t P,
                            uint16_t* __restrict__ dE, float* __restrict__ lossOut, const uint32_t* __restrict__ step,
                            float* __restrict__ lossHist) {
  uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const float* ea = e + (2 * p) * D3;
  const float* eb = e + (2 * p + 1) * D3;
  float diff[D3];
  float d = 0.f;
#pragma unroll
  for (int k = 0; k < D3; ++k) { diff[k] = ea[k] - eb[k]; d += diff[k] * diff[k]; }
  const float y = fminf((float)labels[p], kHCap);
  const float invP = 1.0f / (float)P;
  const float g = 2.f * (d - y) * invP;
#pragma unroll
  for (int k = 0; k < D3; ++k) {
    const uint16_t v = f2bf(2.f * g * diff[k]);
    dE[(2 * p) * D3 + k] = v;
    dE[(2 * p + 1) * D3 + k] = v ^ 0x8000u;     // bf16(-x) = -bf16(x)
  }
  const float l = (d - y) * (d - y) * invP;
  if (lossOut) atomicAdd(lossOut, l);
  if (lossHist) atomicAdd(&lossHist[*step], l);
}

__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, uint16_t* __restrict__ wb, const uint32_t* __restrict__ step, float lr) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N_PARAMS) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float t = (float)(*step + 1u);
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi; v[i] = vi;
  const float mh = mi / (1.f - powf(b1, t));
  const float vh = vi / (1.f - powf(b2, t));
  const float wn = w[i] - lr * mh / (sqrtf(vh) + eps);
  w[i] = wn;
  // bf16 copies of the weight matrices (the biases stay fp32)
  if (i < OFF_B1) wb[WB_W1 + i] = f2bf(wn);
  else if (i >= OFF_W2 && i < OFF_B2) wb[WB_W2 + (i - OFF_W2)] = f2bf(wn);
  else if (i >= OFF_W3 && i < OFF_B3) wb[WB_W3 + (i - OFF_W3)] = f2bf(wn);
}

__global__ void k_refresh_wb(const float* __restrict__ w, uint16_t* __restrict__ wb) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < OFF_B1) wb[WB_W1 + i] = f2bf(w[i]);
  else if (i >= OFF_W2 && i < OFF_B2) wb[WB_W2 + (i - OFF_W2)] = f2bf(w[i]);
  else if (i >= OFF_W3 && i < OFF_B3) wb[WB_W3 + (i - OFF_W3)] = f2bf(w[i]);
}

__global__ void k_set_u32(uint32_t* d, uint32_t v) { *d = v; }
__global__ void k_inc_u32(uint32_t* d) { *d += 1u; }
__global__ void k_zero_f32(float* d, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = 0.f;
}

__device__ __forceinline__ uint64_t sm_next(uint64_t x) { return mix64(x + 0x9E3779B97F4A7C15ull); }

// Near-pair ordering key: popcount in the high bits, then 22 sampled bits.
__global__ void k_pair_keys(const uint32_t* __restrict__ sig, uint64_t n, const uint32_t* __restrict__ bitpos,
                            uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t* s = sig + i * 16;
  uint32_t pc = 0;
#pragma unroll
  for (int w = 0; w < 16; ++w) pc += __popc(s[w]);
  uint32_t k = pc << 22;
  for (int b = 0; b < 22; ++b) {
    const uint32_t q = bitpos[b];
    k |= ((s[q >> 5] >> (q & 31)) & 1u) << b;
  }
  keys[i] = k;
  vals[i] = (uint32_t)i;
}

__global__ void k_inverse_perm(const uint32_t* __restrict__ order, uint64_t n, uint32_t* __restrict__ pos) {
  const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) pos[order[q]] = (uint32_t)q;
}

// Batch of P pairs: near_pct percent are neighbours (+-1..8) in the near-pair
// ordering, the rest random.  Rows 2p, 2p+1 get the two signatures; label =
// Hamming distance.
__global__ void k_sample_batch(const uint32_t* __restrict__ sig, uint64_t n, const uint32_t* __restrict__ order,
                               const uint32_t* __restrict__ pos, uint64_t seed, const uint32_t* __restrict__ step,
                               uint64_t P, uint32_t nearPct, uint32_t* __restrict__ X, uint32_t* __restrict__ labels) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const uint64_t gidx = (uint64_t)(*step) * P + p;
  const uint64_t z1 = mix64(seed ^ ((gidx + 1) * 0x9E3779B97F4A7C15ull));
  const uint64_t z2 = sm_next(z1);
  const uint64_t a = z1 % n;
  uint64_t b;
  if ((p * 37u) % 100u >= nearPct) {
    b = z2 % n;
  } else {
    int64_t d = (int64_t)(z2 % 16) - 8;
    if (d >= 0) d += 1;
    int64_t q = (int64_t)pos[a] + d;
    q = q < 0 ? 0 : (q >= (int64_t)n ? (int64_t)n - 1 : q);
    b = order[q];
  }
  const uint4* sa = reinterpret_cast<const uint4*>(sig + a * 16);
  const uint4* sb = reinterpret_cast<const uint4*>(sig + b * 16);
  uint4* xa = reinterpret_cast<uint4*>(X + (2 * p) * 16);
  uint4* xb = reinterpret_cast<uint4*>(X + (2 * p + 1) * 16);
  uint32_t h = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const uint4 u = sa[q], w = sb[q];
    xa[q] = u; xb[q] = w;
    h += __popc(u.x ^ w.x) + __popc(u.y ^ w.y) + __popc(u.z ^ w.z) + __popc(u.w ^ w.w);
  }
  labels[p] = h;
}

// Working buffers of one forward/backward over R = 2P rows.
struct Work {
  uint64_t R = 0;
  uint16_t *h1 = nullptr, *h2 = nullptr, *dE = nullptr, *dz2 = nullptr, *dz1 = nullptr;
  float* e = nullptr;
  float* part = nullptr;   // split-K partial weight gradients
  void alloc(svd_ctx* c, uint64_t rows) {
    R = rows;
    h1 = dalloc<uint16_t>(c, R * D1); h2 = dalloc<uint16_t>(c, R * D2);
    dE = dalloc<uint16_t>(c, R * D3); dz2 = dalloc<uint16_t>(c, R * D2); dz1 = dalloc<uint16_t>(c, R * D1);
    e = dalloc<float>(c, R * D3);
    const uint64_t np = std::max(split_part_floats(D0, D1, (int)R),
                                 std::max(split_part_floats(D1, D2, (int)R), split_part_floats(D2, D3, (int)R)));
    part = dalloc<float>(c, np);
  }
  void release(svd_ctx* c) {
    dfree(c, h1); dfree(c, h2); dfree(c, dE); dfree(c, dz2); dfree(c, dz1); dfree(c, e); dfree(c, part);
  }
};

void forward(svd_mlp* m, const uint32_t* X, uint64_t rows, uint16_t* h1, uint16_t* h2, float* e) {
  svd_ctx* c = m->ctx;
  GemmArgs g = {};
  g.mValid = rows;
  g.abits = X; g.b = m->wb + WB_W1; g.b_sk = D1; g.b_sn = 1; g.bias = m->params + OFF_B1; g.outb = h1; g.o_sm = D1;
  gemm<A_BITS, E_BIAS_RELU_BF16>(c, rows, D1, D0, g);
  g = {};
  g.mValid = rows;
  g.a = h1; g.a_sm = D1; g.a_sk = 1; g.b = m->wb + WB_W2; g.b_sk = D2; g.b_sn = 1; g.bias = m->params + OFF_B2;
  g.outb = h2; g.o_sm = D2;
  gemm<A_BF16, E_BIAS_RELU_BF16>(c, rows, D2, D1, g);
  g = {};
  g.mValid = rows;
  g.a = h2; g.a_sm = D2; g.a_sk = 1; g.b = m->wb + WB_W3; g.b_sk = D3; g.b_sn = 1; g.bias = m->params + OFF_B3;
  g.outf = e; g.o_sm = D3;
  gemm<A_BF16, E_BIAS_F32>(c, rows, D3, D2, g);
}

// Backward from dE (bf16, R x 32) into m->grads.
void backward(svd_mlp* m, const uint32_t* X, const Work& w) {
  svd_ctx* c = m->ctx;
  const uint64_t R = w.R;
  GemmArgs g = {};
  // dW3 = h2^T dE  (128 x 32), K = R
  g.mValid = D2; g.a = w.h2; g.a_sm = 1; g.a_sk = D2; g.b = w.dE; g.b_sk = D3; g.b_sn = 1;
  g.outf = m->grads + OFF_W3; g.o_sm = D3;
  gemm_wgrad<A_BF16>(c, D2, D3, (int)R, g, w.part);
  k_colsum_bf16<<<D3, 256, 0, c->stream>>>(w.dE, R, D3, m->grads + OFF_B3);
  // dz2 = (dE W3^T) * (h2 > 0)  (R x 128), K = 32
  g = {};
  g.mValid = R; g.a = w.dE; g.a_sm = D3; g.a_sk = 1; g.b = m->wb + WB_W3; g.b_sk = 1; g.b_sn = D3;
  g.mask = w.h2; g.mask_sm = D2; g.outb = w.dz2; g.o_sm = D2;
  gemm<A_BF16, E_MASK_BF16>(c, R, D2, D3, g);
  // dW2 = h1^T dz2  (256 x 128), K = R
  g = {};
  g.mValid = D1; g.a = w.h1; g.a_sm = 1; g.a_sk = D1; g.b = w.dz2; g.b_sk = D2; g.b_sn = 1;
  g.outf = m->grads + OFF_W2; g.o_sm = D2;
  gemm_wgrad<A_BF16>(c, D1, D2, (int)R, g, w.part);
  k_colsum_bf16<<<D2, 256, 0, c->stream>>>(w.dz2, R, D2, m->grads + OFF_B2);
  // dz1 = (dz2 W2^T) * (h1 > 0)  (R x 256), K = 128
  g = {};
  g.mValid = R; g.a = w.dz2; g.a_sm = D2; g.a_sk = 1; g.b = m->wb + WB_W2; g.b_sk = 1; g.b_sn = D2;
  g.mask = w.h1; g.mask_sm = D1; g.outb = w.dz1; g.o_sm = D1;
  gemm<A_BF16, E_MASK_BF16>(c, R, D1, D2, g);
  // dW1 = x^T dz1  (512 x 256), K = R
  g = {};
  g.mValid = D0; g.abits = X; g.b = w.dz1; g.b_sk = D1; g.b_sn = 1; g.outf = m->grads + OFF_W1; g.o_sm = D1;
  gemm_wgrad<A_BITS_T>(c, D0, D1, (int)R, g, w.part);
  k_colsum_bf16<<<D1, 256, 0, c->stream>>>(w.dz1, R, D1, m->grads + OFF_B1);
}

}  // namespace

// Weights: Glorot uniform from a SplitMix64 stream (w = (2u - 1) a,
// u = (z >> 40) 2^-24, a = sqrt(6 / (fan_in + fan_out)) in fp32), zero biases.
void mlp_init(svd_mlp* m, uint64_t seed) {
  svd_ctx* c = m->ctx;
  std::vector<float> h(N_PARAMS, 0.f);
  const int dims[4] = {D0, D1, D2, D3};
  uint64_t state = seed, o = 0;
  for (int l = 0; l < 3; ++l) {
    const uint64_t fi = dims[l], fo = dims[l + 1], n = fi * fo;
    const float a = std::sqrt(6.0f / (float)(fi + fo));
    for (uint64_t i = 1; i <= n; ++i) {
      uint64_t z = state + i * 0x9E3779B97F4A7C15ull;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      z ^= z >> 31;
      const float u = (float)(z >> 40) * (1.0f / 16777216.0f);
      h[o + i - 1] = (u * 2.0f - 1.0f) * a;
    }
    state += n * 0x9E3779B97F4A7C15ull;
    o += n + fo;
  }
  SVD_HIP(hipMemcpyAsync(m->params, h.data(), N_PARAMS * 4, hipMemcpyHostToDevice, c->stream));
  k_refresh_wb<<<div_up(N_PARAMS, 256), 256, 0, c->stream>>>(m->params, m->wb);
  k_zero_f32<<<div_up(N_PARAMS, 256), 256, 0, c->stream>>>(m->m, N_PARAMS);
  k_zero_f32<<<div_up(N_PARAMS, 256), 256, 0, c->stream>>>(m->v, N_PARAMS);
  k_zero_f32<<<div_up(N_PARAMS, 256), 256, 0, c->stream>>>(m->grads, N_PARAMS);
  check_launch("mlp init");
  SVD_HIP(hipStreamSynchronize(c->stream));
  m->steps = 0;
}

void mlp_alloc(svd_mlp* m) {
  svd_ctx* c = m->ctx;
  m->params = dalloc<float>(c, N_PARAMS);
  m->grads = dalloc<float>(c, N_PARAMS);
  m->m = dalloc<float>(c, N_PARAMS);
  m->v = dalloc<float>(c, N_PARAMS);
  m->wb = dalloc<uint16_t>(c, N_WB);
}

void mlp_release(svd_mlp* m) {
  svd_ctx* c = m->ctx;
  dfree(c, m->params); dfree(c, m->grads); dfree(c, m->m); dfree(c, m->v); dfree(c, m->wb);
}

void mlp_set_params(svd_mlp* m, const float* host) {
  svd_ctx* c = m->ctx;
  SVD_HIP(hipMemcpyAsync(m->params, host, N_PARAMS * 4, hipMemcpyHostToDevice, c->stream));
  k_refresh_wb<<<div_up(N_PARAMS, 256), 256, 0, c->stream>>>(m->params, m->wb);
  check_launch("k_refresh_wb");
  SVD_HIP(hipStreamSynchronize(c->stream));
}

void mlp_get_params(svd_mlp* m, float* host) {
  svd_ctx* c = m->ctx;
  SVD_HIP(hipMemcpyAsync(host, m->params, N_PARAMS * 4, hipMemcpyDeviceToHost, c->stream));
  SVD_HIP(hipStreamSynchronize(c->stream));
}

void mlp_embed(svd_mlp* m, const uint32_t* sig, uint64_t n, float* emb) {
  svd_ctx* c = m->ctx;
  const uint64_t chunk = 1ull << 18;
  const uint64_t rows = n < chunk ? n : chunk;
  uint16_t* h1 = dalloc<uint16_t>(c, (rows + 32) * D1);
  uint16_t* h2 = dalloc<uint16_t>(c, (rows + 32) * D2);
  for (uint64_t r0 = 0; r0 < n; r0 += chunk) {
    const uint64_t r = n - r0 < chunk ? n - r0 : chunk;
    forward(m, sig + r0 * 16, r, h1, h2, emb + r0 * D3);
  }
  check_launch("mlp forward");
  dfree(c, h1); dfree(c, h2);
}

void mlp_loss_grad(svd_mlp* m, const uint32_t* X, const uint32_t* labels, uint64_t P, float* lossHost,
                   float* gradsHost) {
  svd_ctx* c = m->ctx;
  Work w;
  w.alloc(c, 2 * P);
  float* loss = dalloc<float>(c, 1);
  k_zero_f32<<<1, 32, 0, c->stream>>>(loss, 1);
  forward(m, X, 2 * P, w.h1, w.h2, w.e);
  k_pair_loss<<<div_up(P, 256), 256, 0, c->stream>>>(w.e, labels, P, w.dE, loss, nullptr, nullptr);
  backward(m, X, w);
  check_launch("mlp loss/grad");
  SVD_HIP(hipMemcpyAsync(lossHost, loss, 4, hipMemcpyDeviceToHost, c->stream));
  SVD_HIP(hipMemcpyAsync(gradsHost, m->grads, N_PARAMS * 4, hipMemcpyDeviceToHost, c->stream));
  SVD_HIP(hipStreamSynchronize(c->stream));
  dfree(c, loss);
  w.release(c);
}

// Pairs are drawn from the nodes of DAG levels lev_lo..lev_hi (signatures
// concatenated).  One step = sample -> forward -> loss -> backward -> Adam ->
// step counter; it is captured once into a HIP graph and replayed.
void mlp_train(svd_mlp* m, svd_octree* o, const svd_mlp_train_params* p, svd_mlp_train_stats* st) {
  svd_ctx* c = m->ctx;
  if (o->state != SVD_S_DAG && o->state != SVD_S_SDAG) throw Error{SVD_ERR_STATE, "training needs the DAG"};
  if (p->lev_lo > p->lev_hi || p->lev_hi + 1 >= o->levels) throw Error{SVD_ERR_ARG, "training levels out of range"};
  if (p->pairs_per_step < 16 || p->pairs_per_step % 16 || p->steps < 1) throw Error{SVD_ERR_ARG, "pairs_per_step must be a multiple of 16, steps >= 1"};
  uint64_t n = 0;
  for (uint32_t l = p->lev_lo; l <= p->lev_hi; ++l) n += o->dag[l].n;
  if (n < 2) throw Error{SVD_ERR_ARG, "not enough nodes to sample pairs from"};
  if (n >= 0xFFFFFFFFull) throw Error{SVD_ERR_CAPACITY, "too many nodes"};
  const uint64_t P = p->pairs_per_step, R = 2 * P;
  const uint32_t nearPct = p->near_pct ? (p->near_pct > 100 ? 100 : p->near_pct) : 50;
  uint32_t* sig = dalloc<uint32_t>(c, n * 16);
  {
    uint64_t off = 0;
    for (uint32_t l = p->lev_lo; l <= p->lev_hi; ++l) {
      lossy_signatures(o, l, sig + off * 16);
      off += o->dag[l].n;
    }
  }
  // near-pair ordering
  uint32_t* keys = dalloc<uint32_t>(c, n);
  uint32_t* vals = dalloc<uint32_t>(c, n);
  uint32_t* keys2 = dalloc<uint32_t>(c, n);
  uint32_t* order = dalloc<uint32_t>(c, n);
  uint32_t* pos = dalloc<uint32_t>(c, n);
  uint32_t* bitpos = dalloc<uint32_t>(c, 22);
  {
    uint32_t hb[22];
    uint64_t s = p->seed ^ 0x5851F42D4C957F2Dull;
    for (int b = 0; b < 22; ++b) {
      s += 0x9E3779B97F4A7C15ull;
      uint64_t z = s;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      z ^= z >> 31;
      hb[b] = (uint32_t)(z % 512);
    }
    SVD_HIP(hipMemcpyAsync(bitpos, hb, sizeof(hb), hipMemcpyHostToDevice, c->stream));
    k_pair_keys<<<div_up(n, 256), 256, 0, c->stream>>>(sig, n, bitpos, keys, vals);
    size_t bytes = 0;
    SVD_HIP(rocprim::radix_sort_pairs(nullptr, bytes, keys, keys2, vals, order, (size_t)n, 0, 32, c->stream));
    void* tmp = c->arena->get(bytes + 16);
    SVD_HIP(rocprim::radix_sort_pairs(tmp, bytes, keys, keys2, vals, order, (size_t)n, 0, 32, c->stream));
    c->arena->put(tmp);
    k_inverse_perm<<<div_up(n, 256), 256, 0, c->stream>>>(order, n, pos);
    check_launch("pair ordering");
  }
  Work w;
  w.alloc(c, R);
  uint32_t* X = dalloc<uint32_t>(c, R * 16);
  uint32_t* labels = dalloc<uint32_t>(c, P);
  uint32_t* dstep = dalloc<uint32_t>(c, 1);
  float* lossHist = dalloc<float>(c, p->steps);
  k_zero_f32<<<div_up(p->steps, 256), 256, 0, c->stream>>>(lossHist, p->steps);
  k_set_u32<<<1, 1, 0, c->stream>>>(dstep, 0u);
  check_launch("train setup");
  const uint32_t step0 = (uint32_t)m->steps;
  // Adam's step index continues across calls: the counter starts at step0
  // for the bias correction, the loss history at 0.
  uint32_t* dadam = dalloc<uint32_t>(c, 1);
  k_set_u32<<<1, 1, 0, c->stream>>>(dadam, step0);
  SVD_HIP(hipStreamSynchronize(c->stream));

  auto enqueue_step = [&]() {
    k_sample_batch<<<div_up(P, 256), 256, 0, c->stream>>>(sig, n, order, pos, p->seed, dstep, P, nearPct, X, labels);
    forward(m, X, R, w.h1, w.h2, w.e);
    k_pair_loss<<<div_up(P, 256), 256, 0, c->stream>>>(w.e, labels, P, w.dE, nullptr, dstep, lossHist);
    backward(m, X, w);
    k_adam<<<div_up(N_PARAMS, 256), 256, 0, c->stream>>>(m->params, m->grads, m->m, m->v, m->wb, dadam, p->lr);
    k_inc_u32<<<1, 1, 0, c->stream>>>(dstep);
    k_inc_u32<<<1, 1, 0, c->stream>>>(dadam);
  };
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  SVD_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  enqueue_step();
  SVD_HIP(hipStreamEndCapture(c->stream, &graph));
  SVD_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  {
    Timer t(c, "mlp_train");
    for (uint32_t s = 0; s < p->steps; ++s) SVD_HIP(hipGraphLaunch(exec, c->stream));
  }
  check_launch("train steps");
  SVD_HIP(hipGraphExecDestroy(exec));
  SVD_HIP(hipGraphDestroy(graph));
  std::vector<float> hist(p->steps);
  SVD_HIP(hipMemcpyAsync(hist.data(), lossHist, p->steps * 4, hipMemcpyDeviceToHost, c->stream));
  SVD_HIP(hipStreamSynchronize(c->stream));
  m->steps += p->steps;
  if (st) {
    const uint32_t q = p->steps < 16 ? p->steps : 16;
    double a = 0, b = 0;
    for (uint32_t i = 0; i < q; ++i) { a += hist[i]; b += hist[p->steps - q + i]; }
    st->first_loss = (float)(a / q);
    st->last_loss = (float)(b / q);
    st->ms_train = c->kernel_ms["mlp_train"];
    st->nodes = n;
    st->pairs = P * p->steps;
  }
  w.release(c);
  dfree(c, X); dfree(c, labels); dfree(c, dstep); dfree(c, dadam); dfree(c, lossHist);
  dfree(c, sig); dfree(c, keys); dfree(c, vals); dfree(c, keys2); dfree(c, order); dfree(c, pos); dfree(c, bitpos);
}

}  // namespace svd
