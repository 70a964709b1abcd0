// Learned candidate encoder for the lossy DAG: MLP 512 -> 256 -> 128 -> 32.
// bf16 weights and activations, fp32 accumulation on the CDNA matrix cores
// (v_mfma_f32_32x32x16_bf16 on LDS-staged tiles; the weight gradients split
// over the batch and summed in a fixed order by the optimizer), hand-written
// forward and backward passes, Adam.
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


constexpr int D0 = 512, D1 = 256, D2 = 128, D3 = 32;
constexpr uint64_t OFF_W1 = 0, OFF_B1 = OFF_W1 + D0 * D1, OFF_W2 = OFF_B1 + D1, OFF_B2 = OFF_W2 + D1 * D2,
                   OFF_W3 = OFF_B2 + D2, OFF_B3 = OFF_W3 + D2 * D3, N_PARAMS = OFF_B3 + D3;
// bf16 weights: W1 | W2 | W3 row-major [in][out], then the same transposed [out][in]
constexpr uint64_t WB_W1 = 0, WB_W2 = WB_W1 + D0 * D1, WB_W3 = WB_W2 + D1 * D2, N_WB = WB_W3 + D2 * D3;
constexpr float kHCap = 32.f;
static_assert(N_PARAMS == SVD_MLP_PARAMS, "parameter layout");

__device__ __forceinline__ uint16_t f2bf(float f) {   // round to nearest even (finite values)
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// C = A (M x K) B (K x N) with bf16 operands and fp32 accumulation.
// Operand layouts:
//   AL_BITS   A[m][k] = bit k of signature m (16 words per signature), as 0/1
//   AL_BITS_T A[m][k] = bit m of signature k
//   AL_MK     A[m][k] = a[m lda + k]        AL_KM  A[m][k] = a[k lda + m]
//   BL_NK     B[k][n] = b[n ldb + k]        BL_KN  B[k][n] = b[k ldb + n]
// Epilogues: bf16(relu(C + bias)); C + bias (fp32); bf16(C) where mask > 0
// (else 0); E_PART: the split's partial C (fp32) and, with partBias, the
// split's column sums of B (a bias gradient when B is a layer's delta);
// E_PAIR_LOSS: e = C + bias is not stored, rows (2p, 2p+1) are pair p and the
// pair loss's gradient w.r.t. e goes out as bf16, its loss into *loss.
enum ALay { AL_BITS = 0, AL_BITS_T = 1, AL_MK = 2, AL_KM = 3 };
enum BLay { BL_NK = 0, BL_KN = 1 };
enum Epi { E_BIAS_RELU_BF16 = 0, E_BIAS_F32 = 1, E_MASK_BF16 = 2, E_PART = 3, E_PAIR_LOSS = 4 };

struct MmArgs {
  uint64_t M;
  int N, K;
  int kPer;                        // E_PART: split z = blockIdx.z covers k in [z kPer, min(K, (z+1) kPer))
  const uint32_t* abits;
  const uint16_t* a; int64_t lda;
  const uint16_t* b; int64_t ldb;
  const float* bias;
  const uint16_t* mask; int64_t ldmask;
  uint16_t* outb; float* outf; int64_t ldo;
  float* part; float* partBias;    // E_PART: part + z M N (row-major, ld N), partBias + z N
  const uint32_t* labels;          // E_PAIR_LOSS: Hamming distance of pair p
  float* loss; const uint32_t* lossStep;   // loss[*lossStep] (loss[0] without lossStep) += the batch loss
  uint32_t* bump;                  // if set, workgroup 0 adds 1 to *bump
};

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

// bf16 0/1 pair from the two low bits of w (element 0 in the low half)
__device__ __forceinline__ uint32_t bit_pair(uint32_t w) {
  return ((w & 1u) ? 0x3F80u : 0u) | ((w & 2u) ? 0x3F800000u : 0u);
}

// Four wavefronts per workgroup, each a 32x32 sub-tile of a TM x TN tile
// (64x64 for WN = 2, 128x32 for WN = 1).  K runs in stages of KB: a stage is
// loaded from global memory into registers while the previous one feeds the
// matrix cores from LDS (two LDS buffers, one barrier per stage).  In LDS, A
// is kept [m][k] and B [n][k], rows padded by 16 bytes; bits are expanded to
// bf16 and the m- or n-contiguous layouts transposed as they are stored.  Per
// 16 of K, lane l of a wavefront reads A[row l%32][k + 8(l/32) + i] and
// B[k + 8(l/32) + i][col l%32], i = 0..7 (16 bytes each), for one
// v_mfma_f32_32x32x16_bf16; accumulator element v of lane l is
// C[8(v/4) + 4(l/32) + v%4][l%32] of the wavefront's sub-tile.
template <int AL, int BL, int EP, int WN, int KB>
__global__ void __launch_bounds__(256) k_mm(MmArgs g) {
  constexpr int WM = 4 / WN, TM = 32 * WM, TN = 32 * WN, ROW = KB + 8;
  constexpr bool kBits = AL == AL_BITS || AL == AL_BITS_T;
  constexpr int A_CH = kBits ? TM * KB / 32 : TM * KB / 8;   // words or 16-byte chunks per stage
  constexpr int B_CH = TN * KB / 8;
  constexpr int NA = (A_CH + 255) / 256, NB = (B_CH + 255) / 256;
  constexpr int QN = 256 / TN, KQ = KB / QN;                   // column-sum split of a stage
  static_assert(EP != E_PART || KQ % 8 == 0, "column sums read 16-byte chunks");
  static_assert(EP != E_PAIR_LOSS || WN == 1, "the pair loss needs all 32 columns of e in one tile");
  __shared__ __attribute__((aligned(16))) uint16_t As[2][TM * ROW];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[2][TN * ROW];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WM, wn = wave / WM;
  const uint64_t m0 = (uint64_t)blockIdx.x * TM;
  const int n0 = blockIdx.y * TN;
  const int z = EP == E_PART ? (int)blockIdx.z : 0;
  const int kBeg = EP == E_PART ? z * g.kPer : 0;
  const int kEnd = EP == E_PART ? min(g.K, kBeg + g.kPer) : g.K;
  const bool colSums = EP == E_PART && g.partBias && blockIdx.x == 0;
  if (g.bump && tid == 0 && blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0) *g.bump += 1u;
  uint4 ra[kBits ? 1 : NA];
  uint32_t rw[kBits ? NA : 1];
  uint4 rb[NB];

  auto load = [&](int k0) {
#pragma unroll
    for (int j = 0; j < NA; ++j) {
      const int c = tid + 256 * j;
      if (c >= A_CH) break;
      if (AL == AL_BITS) {
        const int r = c / (KB / 32), k = k0 + 32 * (c % (KB / 32));
        const uint64_t m = m0 + (uint64_t)r;
        rw[j] = (m < g.M && k < kEnd) ? g.abits[m * 16 + (k >> 5)] : 0u;
      } else if (AL == AL_BITS_T) {
        const int k = k0 + c / (TM / 32);
        const uint64_t m = m0 + 32 * (uint64_t)(c % (TM / 32));
        rw[j] = (m < g.M && k < kEnd) ? g.abits[(uint64_t)k * 16 + (m >> 5)] : 0u;
      } else if (AL == AL_MK) {
        const int r = c / (KB / 8), k = k0 + 8 * (c % (KB / 8));
        const uint64_t m = m0 + (uint64_t)r;
        ra[j] = (m < g.M && k < kEnd) ? *reinterpret_cast<const uint4*>(g.a + (int64_t)m * g.lda + k)
                                      : make_uint4(0u, 0u, 0u, 0u);
      } else {
        const int k = k0 + c / (TM / 8);
        const uint64_t m = m0 + 8 * (uint64_t)(c % (TM / 8));
        ra[j] = (m < g.M && k < kEnd) ? *reinterpret_cast<const uint4*>(g.a + (int64_t)k * g.lda + (int64_t)m)
                                      : make_uint4(0u, 0u, 0u, 0u);
      }
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      const int c = tid + 256 * j;
      if (c >= B_CH) break;
      if (BL == BL_NK) {
        const int nn = c / (KB / 8), k = k0 + 8 * (c % (KB / 8));
        rb[j] = k < kEnd ? *reinterpret_cast<const uint4*>(g.b + (int64_t)(n0 + nn) * g.ldb + k)
                         : make_uint4(0u, 0u, 0u, 0u);
      } else {
        const int k = k0 + c / (TN / 8), nc = 8 * (c % (TN / 8));
        rb[j] = k < kEnd ? *reinterpret_cast<const uint4*>(g.b + (int64_t)k * g.ldb + (n0 + nc))
                         : make_uint4(0u, 0u, 0u, 0u);
      }
    }
  };
  auto store = [&](int buf) {
    uint16_t* as = As[buf];
    uint16_t* bs = Bs[buf];
#pragma unroll
    for (int j = 0; j < NA; ++j) {
      const int c = tid + 256 * j;
      if (c >= A_CH) break;
      if (AL == AL_BITS) {
        const int r = c / (KB / 32), kc = 32 * (c % (KB / 32));
        const uint32_t w = rw[j];
#pragma unroll
        for (int q = 0; q < 4; ++q)
          *reinterpret_cast<uint4*>(as + r * ROW + kc + 8 * q) =
              make_uint4(bit_pair(w >> (8 * q)), bit_pair(w >> (8 * q + 2)), bit_pair(w >> (8 * q + 4)),
                         bit_pair(w >> (8 * q + 6)));
      } else if (AL == AL_BITS_T) {
        const int kr = c / (TM / 32), mc = 32 * (c % (TM / 32));
        const uint32_t w = rw[j];
#pragma unroll
        for (int i = 0; i < 32; ++i) as[(mc + i) * ROW + kr] = ((w >> i) & 1u) ? (uint16_t)0x3F80 : (uint16_t)0;
      } else if (AL == AL_MK) {
        const int r = c / (KB / 8), kc = 8 * (c % (KB / 8));
        *reinterpret_cast<uint4*>(as + r * ROW + kc) = ra[j];
      } else {
        const int kr = c / (TM / 8), mc = 8 * (c % (TM / 8));
        const uint32_t e[4] = {ra[j].x, ra[j].y, ra[j].z, ra[j].w};
#pragma unroll
        for (int i = 0; i < 8; ++i) as[(mc + i) * ROW + kr] = (uint16_t)(e[i >> 1] >> (16 * (i & 1)));
      }
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      const int c = tid + 256 * j;
      if (c >= B_CH) break;
      if (BL == BL_NK) {
        const int nn = c / (KB / 8), kc = 8 * (c % (KB / 8));
        *reinterpret_cast<uint4*>(bs + nn * ROW + kc) = rb[j];
      } else {
        const int kr = c / (TN / 8), nc = 8 * (c % (TN / 8));
        const uint32_t e[4] = {rb[j].x, rb[j].y, rb[j].z, rb[j].w};
#pragma unroll
        for (int i = 0; i < 8; ++i) bs[(nc + i) * ROW + kr] = (uint16_t)(e[i >> 1] >> (16 * (i & 1)));
      }
    }
  };

  // the epilogue's own operands are loaded up front, beside the first stage
  const int nOut = n0 + wn * 32 + (lane & 31);
  const float bn = (EP == E_BIAS_RELU_BF16 || EP == E_BIAS_F32 || EP == E_PAIR_LOSS) ? g.bias[nOut] : 0.f;
  uint32_t keep = 0;   // E_MASK_BF16: bit v = mask of accumulator element v is > 0
  uint32_t lab[8];     // E_PAIR_LOSS: labels of the lane's pairs
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    const uint64_t m = m0 + (uint64_t)(wm * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3));
    if (EP == E_MASK_BF16 && m < g.M) keep |= (bf2f(g.mask[(int64_t)m * g.ldmask + nOut]) > 0.f ? 1u : 0u) << v;
    if (EP == E_PAIR_LOSS && (v & 1) == 0) lab[v >> 1] = m < g.M ? g.labels[m >> 1] : 0u;
  }
  f32x16_t acc;
#pragma unroll
  for (int v = 0; v < 16; ++v) acc[v] = 0.f;
  float colAcc = 0.f;
  const int nk = kEnd > kBeg ? (kEnd - kBeg + KB - 1) / KB : 0;
  if (nk > 0) {
    load(kBeg);
    store(0);
  }
  __syncthreads();
  for (int t = 0; t < nk; ++t) {
    const int buf = t & 1;
    if (t + 1 < nk) load(kBeg + (t + 1) * KB);
    const uint16_t* ar = &As[buf][(wm * 32 + (lane & 31)) * ROW];
    const uint16_t* br = &Bs[buf][(wn * 32 + (lane & 31)) * ROW];
#pragma unroll
    for (int ks = 0; ks < KB; ks += 16) {
      const int kk = ks + 8 * (lane >> 5);
      const bf16x8_t a = *reinterpret_cast<const bf16x8_t*>(ar + kk);
      const bf16x8_t b = *reinterpret_cast<const bf16x8_t*>(br + kk);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    if (EP == E_PART && colSums) {
      const uint16_t* cr = &Bs[buf][(tid % TN) * ROW + (tid / TN) * KQ];
#pragma unroll
      for (int q = 0; q < KQ; q += 8) {
        const uint4 v = *reinterpret_cast<const uint4*>(cr + q);
        const uint32_t e[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int i = 0; i < 4; ++i) colAcc += __uint_as_float(e[i] << 16) + __uint_as_float(e[i] & 0xFFFF0000u);
      }
    }
    if (t + 1 < nk) store(buf ^ 1);
    __syncthreads();
  }
  if (EP == E_PAIR_LOSS) {
    // rows 2p and 2p+1 are accumulator elements v and v+1 (v even) of a lane;
    // d = |e_a - e_b|^2 is summed over the 32 lanes of a half-wavefront
    const int n = nOut;
    const float invP = 1.0f / (float)(g.M / 2);
    float lsum = 0.f;
#pragma unroll
    for (int v = 0; v < 16; v += 2) {
      const uint64_t m = m0 + (uint64_t)(wm * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3));
      const float diff = (acc[v] + bn) - (acc[v + 1] + bn);
      float d = diff * diff;
#pragma unroll
      for (int o = 16; o >= 1; o >>= 1) d += __shfl_xor(d, o, 64);
      if (m < g.M) {
        const float y = fminf((float)lab[v >> 1], kHCap);
        const float gr = 2.f * (d - y) * invP;
        const uint16_t dv = f2bf(2.f * gr * diff);
        g.outb[(int64_t)m * g.ldo + n] = dv;
        g.outb[(int64_t)(m + 1) * g.ldo + n] = dv ^ 0x8000u;     // bf16(-x) = -bf16(x)
        lsum += (d - y) * (d - y) * invP;
      }
    }
    if (n == 0) atomicAdd(&g.loss[g.lossStep ? *g.lossStep : 0u], lsum);
  }
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    if (EP == E_PAIR_LOSS) break;
    const uint64_t m = m0 + (uint64_t)(wm * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3));
    const int n = nOut;
    if (m >= g.M) continue;
    const float x = acc[v];
    if (EP == E_BIAS_RELU_BF16) {
      const float zz = x + bn;
      g.outb[(int64_t)m * g.ldo + n] = f2bf(zz > 0.f ? zz : 0.f);
    } else if (EP == E_BIAS_F32) {
      g.outf[(int64_t)m * g.ldo + n] = x + bn;
    } else if (EP == E_MASK_BF16) {
      g.outb[(int64_t)m * g.ldo + n] = f2bf(((keep >> v) & 1u) ? x : 0.f);
    } else {
      g.part[((uint64_t)z * g.M + m) * (uint64_t)g.N + n] = x;
    }
  }
  if (EP == E_PART && colSums) {
    float* red = reinterpret_cast<float*>(As[0]);   // the stages are done
    red[tid] = colAcc;
    __syncthreads();
    if (tid < TN) {
      float s = 0.f;
      for (int q = 0; q < QN; ++q) s += red[q * TN + tid];
      g.partBias[(uint64_t)z * g.N + n0 + tid] = s;
    }
  }
}

template <int AL, int BL, int EP, int WN, int KB>
void mm(svd_ctx* c, const MmArgs& g, int splits = 1) {
  constexpr int TM = 128 / WN, TN = 32 * WN;
  if (g.N % TN || (EP != E_PART && g.K % KB) || (EP == E_PART && g.kPer % KB) || splits < 1)
    throw Error{SVD_ERR_ARG, "mlp gemm: unsupported shape"};
  if (g.M == 0) return;
  const dim3 grid((uint32_t)div_up(g.M, (uint64_t)TM), (uint32_t)(g.N / TN), (uint32_t)splits);
  k_mm<AL, BL, EP, WN, KB><<<grid, 256, 0, c->stream>>>(g);
}

// Weight gradients: M x N is small and K (the rows of the batch) long, so K
// is split over about kWgradBlocks workgroups; each split writes its partial
// tile and the optimizer (or the gradient read-out) sums the splits in order
// (deterministic, no atomics).
constexpr int kWgradKB = 128;
constexpr uint64_t kWgradBlocks = 512;
constexpr int kMaxSplits = 32;
struct Split { int S, kPer; };
Split wgrad_split(uint64_t M, int N, int K, int TM, int TN) {
  const uint64_t tiles = div_up(M, (uint64_t)TM) * (uint64_t)(N / TN);
  const int maxS = M * (uint64_t)N <= 16384 ? 2 * kMaxSplits : kMaxSplits;   // small partials: split further
  int S = (int)std::min<uint64_t>(maxS, std::max<uint64_t>(1, kWgradBlocks / tiles));
  S = std::min(S, std::max(1, (K + kWgradKB - 1) / kWgradKB));
  int kPer = (K + S - 1) / S;
  kPer = (kPer + kWgradKB - 1) / kWgradKB * kWgradKB;
  return Split{(K + kPer - 1) / kPer, kPer};
}

// The split partial gradients of the three layers.
struct GradParts {
  const float* w[3];
  const float* b[3];
  uint32_t S[3];
};

__device__ __forceinline__ float grad_sum(const GradParts& gp, uint64_t i) {
  constexpr uint64_t offW[3] = {OFF_W1, OFF_W2, OFF_W3}, offB[3] = {OFF_B1, OFF_B2, OFF_B3};
  constexpr uint64_t dout[3] = {D1, D2, D3}, din[3] = {D0, D1, D2};
  const int l = i < OFF_W2 ? 0 : (i < OFF_W3 ? 1 : 2);
  const bool isB = i >= offB[l];
  const float* p = isB ? gp.b[l] + (i - offB[l]) : gp.w[l] + (i - offW[l]);
  const uint64_t stride = isB ? dout[l] : din[l] * dout[l];
  // splits in order; eight loads in flight at a time
  float s = 0.f;
  uint32_t z = 0;
  for (; z + 8 <= gp.S[l]; z += 8) {
    float x[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) x[q] = p[(z + q) * stride];
#pragma unroll
    for (int q = 0; q < 8; ++q) s += x[q];
  }
  for (; z < gp.S[l]; ++z) s += p[z * stride];
  return s;
}

__global__ void k_grad_sum(GradParts gp, float* __restrict__ grads) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N_PARAMS) grads[i] = grad_sum(gp, i);
}

// bf16 copies of weight parameter i (a no-op for the biases, which stay fp32):
// row-major at wb, transposed at wb + N_WB.
__device__ __forceinline__ void store_wb(uint16_t* __restrict__ wb, uint64_t i, float w) {
  constexpr uint64_t offW[3] = {OFF_W1, OFF_W2, OFF_W3}, offB[3] = {OFF_B1, OFF_B2, OFF_B3};
  constexpr uint64_t offWb[3] = {WB_W1, WB_W2, WB_W3};
  constexpr uint64_t dout[3] = {D1, D2, D3}, din[3] = {D0, D1, D2};
  const int l = i < OFF_W2 ? 0 : (i < OFF_W3 ? 1 : 2);
  if (i >= offB[l]) return;
  const uint64_t j = i - offW[l];
  const uint16_t v = f2bf(w);
  wb[offWb[l] + j] = v;
  wb[N_WB + offWb[l] + (j % dout[l]) * din[l] + j / dout[l]] = v;
}

// Adam on the gradients summed from the split partials, at step
// step0 + *step + 1 (bias correction).
__global__ void k_adam(float* __restrict__ w, GradParts gp, float* __restrict__ m, float* __restrict__ v,
                       uint16_t* __restrict__ wb, const uint32_t* __restrict__ step, uint32_t step0, float lr) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N_PARAMS) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float t = (float)(step0 + *step + 1u);
  const float gi = grad_sum(gp, i);
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi; v[i] = vi;
  const float mh = mi / (1.f - powf(b1, t));
  const float vh = vi / (1.f - powf(b2, t));
  const float wn = w[i] - lr * mh / (sqrtf(vh) + eps);
  w[i] = wn;
  store_wb(wb, i, wn);
}

__global__ void k_refresh_wb(const float* __restrict__ w, uint16_t* __restrict__ wb) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N_PARAMS) store_wb(wb, i, w[i]);
}

__global__ void k_set_u32(uint32_t* d, uint32_t v) { *d = v; }
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
// Hamming distance.  The batch number *step is copied to *cur for the later
// kernels of the step (the first layer's GEMM advances *step).
__global__ void k_sample_batch(const uint32_t* __restrict__ sig, uint64_t n, const uint32_t* __restrict__ order,
                               const uint32_t* __restrict__ pos, uint64_t seed, const uint32_t* __restrict__ step,
                               uint64_t P, uint32_t nearPct, uint32_t* __restrict__ X, uint32_t* __restrict__ labels,
                               uint32_t* __restrict__ cur) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p == 0) *cur = *step;
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
  Split sp[3] = {};            // weight-gradient splits of layers 1..3
  float* partW[3] = {};
  float* partB[3] = {};
  void alloc(svd_ctx* c, uint64_t rows) {
    R = rows;
    h1 = dalloc<uint16_t>(c, R * D1); h2 = dalloc<uint16_t>(c, R * D2);
    dE = dalloc<uint16_t>(c, R * D3); dz2 = dalloc<uint16_t>(c, R * D2); dz1 = dalloc<uint16_t>(c, R * D1);
    const int din[3] = {D0, D1, D2}, dout[3] = {D1, D2, D3};
    for (int l = 0; l < 3; ++l) {
      const int wn = dout[l] % 64 ? 1 : 2;
      sp[l] = wgrad_split(din[l], dout[l], (int)R, 128 / wn, 32 * wn);
      partW[l] = dalloc<float>(c, (uint64_t)sp[l].S * din[l] * dout[l]);
      partB[l] = dalloc<float>(c, (uint64_t)sp[l].S * dout[l]);
    }
  }
  GradParts parts() const {
    GradParts gp;
    for (int l = 0; l < 3; ++l) { gp.w[l] = partW[l]; gp.b[l] = partB[l]; gp.S[l] = (uint32_t)sp[l].S; }
    return gp;
  }
  void release(svd_ctx* c) {
    dfree(c, h1); dfree(c, h2); dfree(c, dE); dfree(c, dz2); dfree(c, dz1);
    for (int l = 0; l < 3; ++l) { dfree(c, partW[l]); dfree(c, partB[l]); }
  }
};

// Layers 1 and 2 over `rows` signatures: h1, h2 (bf16).  The weights are
// read transposed ([out][in], k-contiguous).
void forward12(svd_mlp* m, const uint32_t* X, uint64_t rows, uint16_t* h1, uint16_t* h2, uint32_t* bump = nullptr) {
  svd_ctx* c = m->ctx;
  const uint16_t* wt = m->wb + N_WB;
  MmArgs g = {};
  g.M = rows; g.N = D1; g.K = D0; g.bump = bump;
  g.abits = X; g.b = wt + WB_W1; g.ldb = D0; g.bias = m->params + OFF_B1; g.outb = h1; g.ldo = D1;
  mm<AL_BITS, BL_NK, E_BIAS_RELU_BF16, 2, 128>(c, g);
  g = {};
  g.M = rows; g.N = D2; g.K = D1;
  g.a = h1; g.lda = D1; g.b = wt + WB_W2; g.ldb = D1; g.bias = m->params + OFF_B2; g.outb = h2; g.ldo = D2;
  mm<AL_MK, BL_NK, E_BIAS_RELU_BF16, 2, 128>(c, g);
}

// Forward: the embeddings e (fp32) of `rows` signatures.
void forward(svd_mlp* m, const uint32_t* X, uint64_t rows, uint16_t* h1, uint16_t* h2, float* e) {
  forward12(m, X, rows, h1, h2);
  MmArgs g = {};
  g.M = rows; g.N = D3; g.K = D2;
  g.a = h2; g.lda = D2; g.b = m->wb + N_WB + WB_W3; g.ldb = D2; g.bias = m->params + OFF_B3; g.outf = e; g.ldo = D3;
  mm<AL_MK, BL_NK, E_BIAS_F32, 1, 128>(m->ctx, g);
}

// Forward over the R = 2P rows of a pair batch, ending in the pair loss: dE
// (bf16) into w, the batch loss added to loss[*lossStep] (loss[0] if null).
void forward_loss(svd_mlp* m, const uint32_t* X, const uint32_t* labels, const Work& w, float* loss,
                  const uint32_t* lossStep, uint32_t* bump = nullptr) {
  forward12(m, X, w.R, w.h1, w.h2, bump);
  MmArgs g = {};
  g.M = w.R; g.N = D3; g.K = D2;
  g.a = w.h2; g.lda = D2; g.b = m->wb + N_WB + WB_W3; g.ldb = D2; g.bias = m->params + OFF_B3;
  g.outb = w.dE; g.ldo = D3; g.labels = labels; g.loss = loss; g.lossStep = lossStep;
  mm<AL_MK, BL_NK, E_PAIR_LOSS, 1, 128>(m->ctx, g);
}

// Backward from dE (bf16, R x 32) into the split partial gradients of w.
void backward(svd_mlp* m, const uint32_t* X, const Work& w) {
  svd_ctx* c = m->ctx;
  const uint64_t R = w.R;
  MmArgs g = {};
  // dW3 = h2^T dE (128 x 32), db3 = column sums of dE; K = R
  g.M = D2; g.N = D3; g.K = (int)R; g.kPer = w.sp[2].kPer;
  g.a = w.h2; g.lda = D2; g.b = w.dE; g.ldb = D3; g.part = w.partW[2]; g.partBias = w.partB[2];
  mm<AL_KM, BL_KN, E_PART, 1, kWgradKB>(c, g, w.sp[2].S);
  // dz2 = (dE W3^T) * (h2 > 0) (R x 128), K = 32
  g = {};
  g.M = R; g.N = D2; g.K = D3;
  g.a = w.dE; g.lda = D3; g.b = m->wb + WB_W3; g.ldb = D3; g.mask = w.h2; g.ldmask = D2; g.outb = w.dz2; g.ldo = D2;
  mm<AL_MK, BL_NK, E_MASK_BF16, 2, 32>(c, g);
  // dW2 = h1^T dz2 (256 x 128), db2; K = R
  g = {};
  g.M = D1; g.N = D2; g.K = (int)R; g.kPer = w.sp[1].kPer;
  g.a = w.h1; g.lda = D1; g.b = w.dz2; g.ldb = D2; g.part = w.partW[1]; g.partBias = w.partB[1];
  mm<AL_KM, BL_KN, E_PART, 2, kWgradKB>(c, g, w.sp[1].S);
  // dz1 = (dz2 W2^T) * (h1 > 0) (R x 256), K = 128
  g = {};
  g.M = R; g.N = D1; g.K = D2;
  g.a = w.dz2; g.lda = D2; g.b = m->wb + WB_W2; g.ldb = D2; g.mask = w.h1; g.ldmask = D1; g.outb = w.dz1; g.ldo = D1;
  mm<AL_MK, BL_NK, E_MASK_BF16, 2, 128>(c, g);
  // dW1 = x^T dz1 (512 x 256), db1; K = R
  g = {};
  g.M = D0; g.N = D1; g.K = (int)R; g.kPer = w.sp[0].kPer;
  g.abits = X; g.b = w.dz1; g.ldb = D1; g.part = w.partW[0]; g.partBias = w.partB[0];
  mm<AL_BITS_T, BL_KN, E_PART, 2, kWgradKB>(c, g, w.sp[0].S);
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
  m->wb = dalloc<uint16_t>(c, 2 * N_WB);
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
  forward_loss(m, X, labels, w, loss, nullptr);
  backward(m, X, w);
  k_grad_sum<<<div_up(N_PARAMS, 256), 256, 0, c->stream>>>(w.parts(), m->grads);
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
  uint32_t* dstep = dalloc<uint32_t>(c, 2);   // batch counter, the current step's number
  float* lossHist = dalloc<float>(c, p->steps);
  k_zero_f32<<<div_up(p->steps, 256), 256, 0, c->stream>>>(lossHist, p->steps);
  k_set_u32<<<1, 1, 0, c->stream>>>(dstep, 0u);
  check_launch("train setup");
  // Adam's step index continues across calls (step0 + the batch counter),
  // the loss history starts at 0
  const uint32_t step0 = (uint32_t)m->steps;
  SVD_HIP(hipStreamSynchronize(c->stream));

  auto enqueue_step = [&]() {
    k_sample_batch<<<div_up(P, 256), 256, 0, c->stream>>>(sig, n, order, pos, p->seed, dstep, P, nearPct, X, labels,
                                                          dstep + 1);
    forward_loss(m, X, labels, w, lossHist, dstep + 1, dstep);
    backward(m, X, w);
    k_adam<<<div_up(N_PARAMS, 256), 256, 0, c->stream>>>(m->params, w.parts(), m->m, m->v, m->wb, dstep + 1, step0,
                                                         p->lr);
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
    st->ms_train = timer_ms(c, "mlp_train");
    st->nodes = n;
    st->pairs = P * p->steps;
  }
  w.release(c);
  dfree(c, X); dfree(c, labels); dfree(c, dstep); dfree(c, lossHist);
  dfree(c, sig); dfree(c, keys); dfree(c, vals); dfree(c, keys2); dfree(c, order); dfree(c, pos); dfree(c, bitpos);
}

}  // namespace svd
