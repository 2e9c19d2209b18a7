// MLP cost model kernels: features, MFMA GEMMs (bf16 in, fp32 accumulate),
// forward, backward, Adam.  See mlp.h for the model.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "device_common.h"
#include "mlp.h"

namespace cubff {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__host__ __device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
  return (uint16_t)(u >> 16);
}

__host__ __device__ __forceinline__ float bf2f(uint16_t h) {
  const uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

enum GemmEpi : int {
  kEpiF32 = 0,        // out (fp32) = C
  kEpiBiasRelu = 1,   // out (bf16) = relu(C + bias)
  kEpiBf16 = 2,       // out (bf16) = C
  kEpiMaskBf16 = 3,   // out (bf16) = C * (mask != 0)
  kEpiF32Add = 4,     // out (fp32) += C (atomic; split-K partial sums)
};

// C[M, N] = A[M, K] B[K, N] with A(i, k) = A[i * a_si + k * a_sk] and
// B(k, j) = B[k * b_sk + j * b_sj].  M, N multiples of 64; K of 16.
// Block: 4 waves, 64 x 64 tile; each wave a 32 x 32 tile of four
// v_mfma_f32_16x16x16_bf16 accumulators.  Operand fragments: lane l holds
// A(i0 + l % 16, k0 + 4 (l / 16) + 0..3) and B(k0 + 4 (l / 16) + 0..3,
// j0 + l % 16); the accumulator holds C(i0 + 4 (l / 16) + r, j0 + l % 16).
// AC / BC: the four k values are contiguous (one 8-byte load).  Block z
// covers k in [z * K, (z + 1) * K) (split-K with kEpiF32Add).
template <bool AC, bool BC, int EPI>
__global__ __launch_bounds__(256) void k_gemm(int K, const uint16_t *__restrict__ A,
                                              int64_t a_si, int64_t a_sk,
                                              const uint16_t *__restrict__ B,
                                              int64_t b_sk, int64_t b_sj,
                                              const float *__restrict__ bias,
                                              void *__restrict__ out, int64_t ldo,
                                              const uint16_t *__restrict__ mask) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t i0 = (int64_t)blockIdx.y * 64 + (w >> 1) * 32;
  const int64_t j0 = (int64_t)blockIdx.x * 64 + (w & 1) * 32;
  const int r16 = lane & 15, kq = (lane >> 4) * 4;
  f32x4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int kbeg = (int)blockIdx.z * K;
  for (int k0 = kbeg; k0 < kbeg + K; k0 += 16) {
    bf16x4 fa[2], fb[2];
#pragma unroll
    for (int t = 0; t < 2; t++) {
      const int64_t row = i0 + t * 16 + r16;
      if constexpr (AC) {
        fa[t] = *reinterpret_cast<const bf16x4 *>(A + row * a_si + k0 + kq);
      } else {
#pragma unroll
        for (int q = 0; q < 4; q++) fa[t][q] = (short)A[row * a_si + (int64_t)(k0 + kq + q) * a_sk];
      }
      const int64_t col = j0 + t * 16 + r16;
      if constexpr (BC) {
        fb[t] = *reinterpret_cast<const bf16x4 *>(B + col * b_sj + k0 + kq);
      } else {
#pragma unroll
        for (int q = 0; q < 4; q++) fb[t][q] = (short)B[(int64_t)(k0 + kq + q) * b_sk + col * b_sj];
      }
    }
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++)
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(fa[a], fb[b], acc[a][b], 0, 0, 0);
  }
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int64_t row = i0 + a * 16 + kq + r, col = j0 + b * 16 + r16;
        float v = acc[a][b][r];
        if constexpr (EPI == kEpiF32) {
          reinterpret_cast<float *>(out)[row * ldo + col] = v;
        } else if constexpr (EPI == kEpiF32Add) {
          atomicAdd(reinterpret_cast<float *>(out) + row * ldo + col, v);
        } else {
          if constexpr (EPI == kEpiBiasRelu) v = fmaxf(v + bias[col], 0.f);
          if constexpr (EPI == kEpiMaskBf16) v = mask[row * ldo + col] ? v : 0.f;
          reinterpret_cast<uint16_t *>(out)[row * ldo + col] = f2bf(v);
        }
      }
}

template <bool AC, bool BC, int EPI>
static hipError_t gemm(hipStream_t st, int M, int N, int K, const uint16_t *A, int64_t a_si,
                       int64_t a_sk, const uint16_t *B, int64_t b_sk, int64_t b_sj,
                       const float *bias, void *out, int64_t ldo, const uint16_t *mask,
                       int splits = 1) {
  if (M % 64 || N % 64 || K % (16 * splits)) return hipErrorInvalidValue;
  if (splits > 1 && EPI != kEpiF32Add) return hipErrorInvalidValue;
  if (M == 0) return hipSuccess;
  hipLaunchKernelGGL((k_gemm<AC, BC, EPI>), dim3(N / 64, M / 64, splits), dim3(256), 0, st,
                     K / splits, A, a_si, a_sk, B, b_sk, b_sj, bias, out, ldo, mask);
  return hipGetLastError();
}

// Splits of a weight-gradient GEMM over the batch dimension K (a multiple of
// 64): chunks of at least 256 rows, at most 32 of them.
static int wgrad_splits(uint64_t K) {
  int s = 1;
  while (s < 32 && K % (uint64_t)(128 * s) == 0 && K / (uint64_t)(2 * s) >= 256) s *= 2;
  return s;
}

static inline uint64_t pad64(uint64_t n) { return (n + 63) / 64 * 64; }
static inline uint64_t pad128(uint64_t n) { return (n + 127) / 128 * 128; }

// Forward GEMM out[M, 256] = A[M, K] W with W given as WT[256][K] (both
// contiguous along K), M a multiple of 128.  Block: 8 waves, 128 x 256 tile;
// wave w covers rows 32 (w / 2) .. +32 and columns 128 (w % 2) .. +128 (2 x 8
// MFMA tiles).  The weights are staged through LDS in K-chunks of 64 (rows
// padded to 68 bf16 so the 16 lanes of a fragment read hit distinct banks),
// so every activation row is read from memory once.
template <int K, int EPI>
__global__ __launch_bounds__(512) void k_gemm_fwd(const uint16_t *__restrict__ A,
                                                  const uint16_t *__restrict__ WT,
                                                  const float *__restrict__ bias,
                                                  uint16_t *__restrict__ out) {
  constexpr int KC = 64, KP = KC + 4;
  __shared__ __align__(16) uint16_t bs[kMlpHid * KP];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t i0 = (int64_t)blockIdx.x * 128 + (w >> 1) * 32;
  const int j0 = (w & 1) * 128;
  const int r16 = lane & 15, kq = (lane >> 4) * 4;
  f32x4 acc[2][8];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 8; b++) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int kc = 0; kc < K; kc += KC) {
    __syncthreads();
    // 256 rows x 64 bf16 = 2048 chunks of 8 bytes; 4 per thread.
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const int idx = threadIdx.x + c * 512;
      const int row = idx >> 3, part = idx & 7;  // 8 chunks of 8 bf16 per row
      const uint4 v = *reinterpret_cast<const uint4 *>(WT + (int64_t)row * K + kc + part * 8);
      uint32_t *dst = reinterpret_cast<uint32_t *>(bs + row * KP + part * 8);
      dst[0] = v.x;
      dst[1] = v.y;
      dst[2] = v.z;
      dst[3] = v.w;
    }
    __syncthreads();
    // All A fragments of this K-chunk are loaded up front (8 loads in flight).
    bf16x4 fa_all[KC / 16][2];
#pragma unroll
    for (int q = 0; q < KC / 16; q++)
#pragma unroll
      for (int t = 0; t < 2; t++)
        fa_all[q][t] =
            *reinterpret_cast<const bf16x4 *>(A + (i0 + t * 16 + r16) * K + kc + q * 16 + kq);
#pragma unroll
    for (int ks = 0; ks < KC; ks += 16) {
      const bf16x4 *fa = fa_all[ks / 16];
#pragma unroll
      for (int b = 0; b < 8; b++) {
        const bf16x4 fb = *reinterpret_cast<const bf16x4 *>(bs + (j0 + b * 16 + r16) * KP + ks + kq);
#pragma unroll
        for (int t = 0; t < 2; t++)
          acc[t][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(fa[t], fb, acc[t][b], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 8; b++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int64_t row = i0 + a * 16 + kq + r;
        const int col = j0 + b * 16 + r16;
        float v = acc[a][b][r];
        if constexpr (EPI == kEpiBiasRelu) v = fmaxf(v + bias[col], 0.f);
        out[row * kMlpHid + col] = f2bf(v);
      }
}


// ---------------------------------------------------------------------------
// Features: byte histogram / 128.  16 tapes per block, 16 threads per tape
// (8 bytes each).
__global__ __launch_bounds__(256) void k_features(const uint8_t *__restrict__ tapes,
                                                  const uint32_t *__restrict__ rows,
                                                  uint64_t count, uint16_t *__restrict__ x) {
  __shared__ uint32_t h[16][kMlpIn];
  const int t = threadIdx.x;
  for (int i = t; i < 16 * kMlpIn; i += 256) (&h[0][0])[i] = 0;
  __syncthreads();
  const int ti = t >> 4, c = t & 15;
  const uint64_t r = (uint64_t)blockIdx.x * 16 + ti;
  if (r < count) {
    const uint64_t src = rows ? rows[r] : r;
    const uint8_t *tp = tapes + src * kTape2;
#pragma unroll
    for (int q = 0; q < 8; q++) atomicAdd(&h[ti][tp[c * 8 + q]], 1u);
  }
  __syncthreads();
  for (int i = t; i < 16 * kMlpIn; i += 256) {
    const int tj = i / kMlpIn, f = i % kMlpIn;
    const uint64_t rr = (uint64_t)blockIdx.x * 16 + tj;
    if (rr < count) x[rr * kMlpIn + f] = f2bf((float)h[tj][f] * (1.0f / 128));
  }
}

// y[r] = a2[r] . w3 + b3.
__global__ void k_out(const uint16_t *__restrict__ a2, const float *__restrict__ params,
                      float *__restrict__ y, uint64_t count) {
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= count) return;
  const float *w3 = params + kOffW3;
  const uint4 *row = reinterpret_cast<const uint4 *>(a2 + r * kMlpHid);
  float s = params[kOffB3];
  for (int q = 0; q < kMlpHid / 8; q++) {
    const uint4 v = row[q];
    const uint32_t u[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int e = 0; e < 4; e++) {
      s += bf2f((uint16_t)(u[e] & 0xffff)) * w3[q * 8 + 2 * e];
      s += bf2f((uint16_t)(u[e] >> 16)) * w3[q * 8 + 2 * e + 1];
    }
  }
  y[r] = s;
}

// dy = (y - t) / n (zero for padding rows); d2 = dy * w3 * relu'(a2).  One
// block per row, one thread per column; dy is left in y.
__global__ __launch_bounds__(kMlpHid) void k_out_bwd(
    const uint16_t *__restrict__ a2, const float *__restrict__ params, float *__restrict__ y,
    const float *__restrict__ target, uint16_t *__restrict__ d2, uint64_t count) {
  const uint64_t r = blockIdx.x;
  const int j = threadIdx.x;
  const float dy = r < count ? (y[r] - target[r]) / (float)count : 0.f;
  __syncthreads();  // every thread has read y[r] before it is overwritten
  if (j == 0) y[r] = dy;
  const uint16_t a = a2[r * kMlpHid + j];
  d2[r * kMlpHid + j] = f2bf(a ? dy * params[kOffW3 + j] : 0.f);
}

// dw3[j] += sum over a chunk of rows of dy[r] a2[r][j]; db3 likewise; and
// the loss 0.5 mean((y - t)^2) = sum 0.5 n dy^2.  Grid: one block of 256
// threads (one per j) per kRedRows rows; results are accumulated with atomics
// into zeroed gradients.
constexpr int kRedRows = 32;
__global__ void k_out_wgrad(const uint16_t *__restrict__ a2, const float *__restrict__ dy,
                            uint64_t count, float *__restrict__ grads,
                            float *__restrict__ loss) {
  const int j = threadIdx.x;
  const uint64_t r0 = (uint64_t)blockIdx.x * kRedRows;
  const uint64_t r1 = min(count, r0 + kRedRows);
  float s = 0.f, sb = 0.f, sl = 0.f;
  for (uint64_t r = r0; r < r1; r++) {
    s += dy[r] * bf2f(a2[r * kMlpHid + j]);
    sb += dy[r];
    sl += 0.5f * (float)count * dy[r] * dy[r];
  }
  atomicAdd(&grads[kOffW3 + j], s);
  if (j == 0) {
    atomicAdd(&grads[kOffB3], sb);
    atomicAdd(loss, sl);
  }
}

// db[j] += sum over a chunk of rows of d[r][j].
__global__ void k_bias_grad(const uint16_t *__restrict__ d, uint64_t count,
                            float *__restrict__ db) {
  const int j = threadIdx.x;
  const uint64_t r0 = (uint64_t)blockIdx.x * kRedRows;
  const uint64_t r1 = min(count, r0 + kRedRows);
  float s = 0.f;
  for (uint64_t r = r0; r < r1; r++) s += bf2f(d[r * kMlpHid + j]);
  atomicAdd(&db[j], s);
}

__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, uint64_t n, float lr, float c1, float c2) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1 - b1) * gi;
  const float vi = b2 * v[i] + (1 - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// bf16 copies of one weight matrix W[in][out]: w (row-major) and wt.
__global__ void k_to_bf16(const float *__restrict__ W, int in, int out, uint16_t *__restrict__ w,
                          uint16_t *__restrict__ wt) {
  const int i = blockIdx.x, j = threadIdx.x;
  if (i >= in || j >= out) return;
  const uint16_t h = f2bf(W[i * out + j]);
  w[i * out + j] = h;
  wt[j * in + i] = h;
}

__global__ void k_targets(const uint32_t *__restrict__ ops, const uint32_t *__restrict__ rows,
                          uint64_t count, float *__restrict__ target) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) target[i] = log2f((float)ops[rows ? rows[i] : i] + 1.0f);
}

// ---------------------------------------------------------------------------
// Fused inference, weight-stationary: one persistent workgroup per CU holds
// both hidden layers' weights in registers for the whole launch and streams
// the tapes through, 64 per pass, so no weight is read from memory more than
// once per workgroup.
//
// Workgroup: 8 waves; wave w owns output columns 32 w .. 32 w + 31 of both
// layers, held as the MFMA's A fragments (the transposed product
// C^T = W^T X^T: 2 column tiles x 8 K-steps of 32, four VGPRs each: 64
// VGPRs per layer).  Fragment layout of v_mfma_f32_16x16x32_bf16: lane l
// holds A(l % 16, 8 (l / 16) + 0..7), B(8 (l / 16) + 0..7, l % 16) and
// C(4 (l / 16) + 0..3, l % 16), so a lane's accumulator is four consecutive
// columns of one row.  A pass:
//   1. byte counts of 64 tapes into an LDS histogram (four u8 counters per
//      dword, LDS integer atomics; counts never exceed 128), from tapes
//      loaded into registers during the previous pass, then converted once,
//      cooperatively, to a bf16 copy (exact: integers up to 256), zeroing
//      the counters for the next pass;
//   2. a1 = relu(counts (E W1 / 128) + b1): B fragments are 16-byte LDS
//      reads of the bf16 counts; a1 written to LDS as bf16, four columns per
//      8-byte store (v_cvt_pk_bf16_f32, round to nearest even);
//   3. a2 = relu(a1 W2 + b2), rounded to bf16 as in the unfused path, and
//      y = a2 . w3 + b3 reduced across lanes and waves in LDS.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kInfRows = 64;       // tapes per pass
constexpr int kInfThreads = 512;
constexpr int kRowStride = 264;    // bf16 per histogram / activation row (256 + pad)

constexpr int kCountStride = 264;  // bytes per u8 count row (256 + pad)

// Workgroup barrier ordering LDS accesses only: unlike __syncthreads() it
// does not wait for outstanding global loads, so the next pass's tapes stay
// in flight across the barriers of the current pass.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Round to bf16 (nearest even) with the hardware conversion.
__device__ __forceinline__ uint16_t to_bf16(float x) {
  return __builtin_bit_cast(uint16_t, (__bf16)x);
}

__global__ __launch_bounds__(kInfThreads) void k_predict(const uint8_t *__restrict__ tapes,
                                                 const uint32_t *__restrict__ halves,
                                                 unsigned int *__restrict__ claim,
                                                 uint64_t count,
                                                 const uint16_t *__restrict__ ew1t,
                                                 const uint16_t *__restrict__ w2t,
                                                 const float *__restrict__ vec,
                                                 float *__restrict__ y) {
  __shared__ __align__(16) uint32_t cnt8[kInfRows * kCountStride / 4];
  __shared__ __align__(16) uint16_t hist[kInfRows * kRowStride];
  __shared__ __align__(16) uint16_t act[kInfRows * kRowStride];
  __shared__ float ypart[kInfRows];
  __shared__ uint32_t claimed;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int r16 = lane & 15, kg = lane >> 4;
  const int col0 = 32 * w;

  // This wave's weights, bias and output weights, for the whole launch.
  bf16x8 wb1[2][8], wb2[2][8];
#pragma unroll
  for (int c = 0; c < 2; c++)
#pragma unroll
    for (int s = 0; s < 8; s++) {
      const int64_t off = (int64_t)(col0 + 16 * c + r16) * kMlpHid + 32 * s + 8 * kg;
      wb1[c][s] = *reinterpret_cast<const bf16x8 *>(ew1t + off);
      wb2[c][s] = *reinterpret_cast<const bf16x8 *>(w2t + off);
    }
  // Weights are the MFMA's A operand and activations its B operand, so the
  // accumulator of tile (t, c) holds, in lane l, row 16t + l % 16 at the
  // four consecutive columns col0 + 16c + 4(l / 16) + r, r = 0..3.
  float b1[2][4], b2[2][4], w3[2][4];
#pragma unroll
  for (int c = 0; c < 2; c++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int col = col0 + 16 * c + 4 * kg + r;
      b1[c][r] = vec[col];
      b2[c][r] = vec[kMlpHid + col];
      w3[c][r] = vec[2 * kMlpHid + col];
    }
  const float b3 = vec[3 * kMlpHid];

  // Thread tid loads 16 bytes (chunk tid % 8) of tape tid / 8 of each pass:
  // bytes 16 (tid % 4) .. +15 of 64-byte piece 2 row + (tid % 8) / 4, which
  // is that piece of the contiguous tapes or the one `halves` names.
  const int tr = tid >> 3, tpart = tid & 7;
  auto piece = [&](uint64_t row) -> const uint4 * {
    const uint64_t h = 2 * row + (tpart >> 2);
    return reinterpret_cast<const uint4 *>(tapes + (halves ? (uint64_t)halves[h] : h) * kTape) +
           (tpart & 3);
  };
  const uint64_t passes = (count + kInfRows - 1) / kInfRows;
  uint4 next = make_uint4(0, 0, 0, 0);
  for (int i = tid; i < kInfRows * kCountStride / 4; i += kInfThreads) cnt8[i] = 0;
  if (tid < kInfRows) ypart[tid] = 0.f;
  lds_barrier();
  // Workgroup b starts with pass b.  Its next pass is b + grid, b + 2 grid,
  // ... or, with `claim` (zero at launch), the next unclaimed one: beside
  // other kernels, workgroups placed late then take fewer passes.
  uint64_t p = blockIdx.x;
  if (p < passes && p * kInfRows + tr < count) next = *piece(p * kInfRows + tr);
  while (p < passes) {
    const uint64_t row0 = p * kInfRows;
    const bool have = row0 + tr < count;
    const uint4 cur = next;
    // 1. Counts into cnt8 (zeroed by the previous pass's conversion, with
    // barriers since; hist, act and ypart are next written after barriers).
    // `claimed` was last read before this pass's earlier barriers.
    uint64_t pn = p + gridDim.x;
    if (claim) {
      if (tid == 0) claimed = gridDim.x + atomicAdd(claim, 1u);
    } else if (pn < passes && pn * kInfRows + tr < count) {
      next = *piece(pn * kInfRows + tr);
    }
    if (have) {
      const uint32_t wd[4] = {cur.x, cur.y, cur.z, cur.w};
      uint32_t *hr = cnt8 + tr * (kCountStride / 4);
#pragma unroll
      for (int b = 0; b < 16; b++) {
        const uint32_t v = (wd[b >> 2] >> (8 * (b & 3))) & 255u;
        atomicAdd(&hr[v >> 2], 1u << (8 * (v & 3)));
      }
    }
    lds_barrier();
    if (claim) {
      pn = claimed;
      if (pn < passes && pn * kInfRows + tr < count) next = *piece(pn * kInfRows + tr);
    }
    // Counts to bf16, once for all waves: thread tid converts 32 counts
    // (row tid / 8, bins 32 (tid % 8) .. +31).
    {
      uint32_t *src = cnt8 + tr * (kCountStride / 4) + 8 * tpart;
      uint32_t *dst = reinterpret_cast<uint32_t *>(hist + tr * kRowStride + 32 * tpart);
      uint32_t cv[8];
#pragma unroll
      for (int q = 0; q < 8; q++) cv[q] = src[q];
#pragma unroll
      for (int q = 0; q < 8; q++) src[q] = 0;  // ready for the next pass
#pragma unroll
      for (int q = 0; q < 8; q++) {
        const uint32_t v = cv[q];
        const __bf16 h0 = (__bf16)(float)(v & 255u), h1 = (__bf16)(float)((v >> 8) & 255u);
        const __bf16 h2 = (__bf16)(float)((v >> 16) & 255u), h3 = (__bf16)(float)(v >> 24);
        dst[2 * q] = (uint32_t)__builtin_bit_cast(uint16_t, h0) |
                     (uint32_t)__builtin_bit_cast(uint16_t, h1) << 16;
        dst[2 * q + 1] = (uint32_t)__builtin_bit_cast(uint16_t, h2) |
                         (uint32_t)__builtin_bit_cast(uint16_t, h3) << 16;
      }
    }
    lds_barrier();

    // 2. a1 = relu(x (E W1) + b1).
    f32x4 acc[4][2];
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
      for (int c = 0; c < 2; c++) acc[t][c] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < 8; s++) {
#pragma unroll
      for (int t = 0; t < 4; t++) {
        const bf16x8 fa = *reinterpret_cast<const bf16x8 *>(hist + (16 * t + r16) * kRowStride +
                                                            32 * s + 8 * kg);
#pragma unroll
        for (int c = 0; c < 2; c++)
          acc[t][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wb1[c][s], fa, acc[t][c], 0, 0, 0);
      }
    }
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
      for (int c = 0; c < 2; c++) {
        uint32_t w[2];
#pragma unroll
        for (int h = 0; h < 2; h++)
          w[h] = (uint32_t)to_bf16(fmaxf(acc[t][c][2 * h] + b1[c][2 * h], 0.f)) |
                 (uint32_t)to_bf16(fmaxf(acc[t][c][2 * h + 1] + b1[c][2 * h + 1], 0.f)) << 16;
        *reinterpret_cast<uint2 *>(act + (16 * t + r16) * kRowStride + col0 + 16 * c + 4 * kg) =
            make_uint2(w[0], w[1]);
        acc[t][c] = f32x4{0.f, 0.f, 0.f, 0.f};
      }
    lds_barrier();

    // 3. a2 = relu(a1 W2 + b2); y = a2 . w3 + b3.
#pragma unroll
    for (int s = 0; s < 8; s++) {
#pragma unroll
      for (int t = 0; t < 4; t++) {
        const bf16x8 fa = *reinterpret_cast<const bf16x8 *>(act + (16 * t + r16) * kRowStride +
                                                            32 * s + 8 * kg);
#pragma unroll
        for (int c = 0; c < 2; c++)
          acc[t][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wb2[c][s], fa, acc[t][c], 0, 0, 0);
      }
    }
    // Row sums of a2 . w3 over this wave's 32 columns: lane l holds, for
    // each t, the sum over its 8 columns of row 16t + l % 16; the four lanes
    // l % 16 + 16 kg of a row add theirs in two exchanges that leave row
    // 16 kg + l % 16 in lane l.
    float v[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
      float x = 0.f;
#pragma unroll
      for (int c = 0; c < 2; c++)
#pragma unroll
        for (int r = 0; r < 4; r++)
          x += (float)(__bf16)fmaxf(acc[t][c][r] + b2[c][r], 0.f) * w3[c][r];
      v[t] = x;
    }
    {
      const bool hi = kg & 2;  // partner: lane ^ 32
#pragma unroll
      for (int j = 0; j < 2; j++) {
        const float keep = hi ? v[2 + j] : v[j];
        const float send = hi ? v[j] : v[2 + j];
        v[j] = keep + __shfl_xor(send, 32, 64);
      }
      const bool lo = kg & 1;  // partner: lane ^ 16
      const float keep = lo ? v[1] : v[0];
      const float send = lo ? v[0] : v[1];
      v[0] = keep + __shfl_xor(send, 16, 64);
    }
    atomicAdd(&ypart[16 * kg + r16], v[0]);
    lds_barrier();
    if (tid < kInfRows) {
      if (row0 + tid < count) y[row0 + tid] = ypart[tid] + b3;
      ypart[tid] = 0.f;  // only this thread reads it; added to after barriers
    }
    p = pn;
  }
}

__global__ void k_zero_claim(unsigned int *claim) { *claim = 0; }

hipError_t mlp_predict(Mlp &m, const uint8_t *tapes, uint64_t count, float *y,
                       hipStream_t st, const uint32_t *halves) {
  if (count == 0) return hipSuccess;
  if (!m.folded_valid) {
    hipError_t e = mlp_fold(m, st);
    if (e != hipSuccess) return e;
  }
  return mlp_predict_live(m, tapes, count, y, st, halves);
}

hipError_t mlp_predict_live(const Mlp &m, const uint8_t *tapes, uint64_t count, float *y,
                            hipStream_t st, const uint32_t *halves) {
  if (count == 0) return hipSuccess;
  int dev = 0, cus = 0;
  hipError_t e;
  if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
  if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess)
    return e;
  const uint64_t passes = (count + kInfRows - 1) / kInfRows;
  const unsigned grid = (unsigned)(passes < (uint64_t)cus ? passes : (uint64_t)cus);
  if (m.claim) hipLaunchKernelGGL(k_zero_claim, dim3(1), dim3(1), 0, st, m.claim);
  hipLaunchKernelGGL(k_predict, dim3(grid), dim3(kInfThreads), 0, st, tapes, halves, m.claim,
                     count, m.live.ew1t,
                     m.live.w2t, m.live.vec, y);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------

template <typename T>
static hipError_t zalloc(T **p, uint64_t n) {
  hipError_t e = hipMalloc(reinterpret_cast<void **>(p), (n ? n : 1) * sizeof(T));
  if (e == hipSuccess) e = hipMemset(*p, 0, (n ? n : 1) * sizeof(T));
  return e;
}

hipError_t mlp_alloc(Mlp &m, uint64_t capacity) {
  mlp_free(m);
  capacity = pad128(capacity ? capacity : 128);
  m.capacity = capacity;
  hipError_t e;
  if ((e = zalloc(&m.params, kNumParams)) || (e = zalloc(&m.grads, kNumParams)) ||
      (e = zalloc(&m.m, kNumParams)) || (e = zalloc(&m.v, kNumParams)) ||
      (e = zalloc(&m.e_w, (uint64_t)kMlpIn * kMlpHid)) ||
      (e = zalloc(&m.e_wt, (uint64_t)kMlpIn * kMlpHid)) ||
      (e = zalloc(&m.w1, (uint64_t)kMlpHid * kMlpHid)) ||
      (e = zalloc(&m.w1t, (uint64_t)kMlpHid * kMlpHid)) ||
      (e = zalloc(&m.w2, (uint64_t)kMlpHid * kMlpHid)) ||
      (e = zalloc(&m.w2t, (uint64_t)kMlpHid * kMlpHid)) ||
      (e = zalloc(&m.live.ew1t, (uint64_t)kMlpIn * kMlpHid)) ||
      (e = zalloc(&m.live.w2t, (uint64_t)kMlpHid * kMlpHid)) ||
      (e = zalloc(&m.live.vec, kInfVec)) ||
      (e = zalloc(&m.next.ew1t, (uint64_t)kMlpIn * kMlpHid)) ||
      (e = zalloc(&m.next.w2t, (uint64_t)kMlpHid * kMlpHid)) ||
      (e = zalloc(&m.next.vec, kInfVec)) ||
      (e = zalloc(&m.x, capacity * kMlpIn)) || (e = zalloc(&m.p, capacity * kMlpHid)) ||
      (e = zalloc(&m.a1, capacity * kMlpHid)) || (e = zalloc(&m.a2, capacity * kMlpHid)) ||
      (e = zalloc(&m.d2, capacity * kMlpHid)) || (e = zalloc(&m.d1, capacity * kMlpHid)) ||
      (e = zalloc(&m.dp, capacity * kMlpHid)) || (e = zalloc(&m.y, capacity)) ||
      (e = zalloc(&m.target, capacity)) || (e = zalloc(&m.loss, 1)) ||
      (e = zalloc(&m.claim, 1)))
    return e;
  return hipSuccess;
}

void mlp_free(Mlp &m) {
  void *ps[] = {m.params, m.grads, m.m, m.v, m.e_w, m.e_wt, m.w1, m.w1t, m.w2, m.w2t,
                m.live.ew1t, m.live.w2t, m.live.vec, m.next.ew1t, m.next.w2t, m.next.vec,
                m.x, m.p, m.a1, m.a2, m.d2, m.d1, m.dp, m.y, m.target, m.loss, m.claim};
  for (void *p : ps)
    if (p) (void)hipFree(p);
  m = Mlp{};
}

// Folded first layer E W1 / 128 (fp32 GEMM on the fp32 masters), stored as
// bf16 [256][256] for the inference kernel, which feeds it raw counts.
__global__ void k_fold(const float *__restrict__ params, uint16_t *__restrict__ ew1t) {
  const int i = blockIdx.x, j = threadIdx.x;  // i: input (256), j: hidden (256)
  const float *E = params + kOffE, *W1 = params + kOffW1;
  float s = 0.f;
  for (int k = 0; k < kMlpHid; k++) s += E[i * kMlpHid + k] * W1[k * kMlpHid + j];
  // Scaled by the feature normalisation 1/128 (exact in bf16), so inference
  // multiplies raw counts.
  ew1t[j * kMlpIn + i] = f2bf(s * (1.0f / 128));
}

// b1, b2, w3, b3 of the fp32 masters into an inference vector.
__global__ void k_inf_vec(const float *__restrict__ params, float *__restrict__ vec) {
  const int j = threadIdx.x;
  vec[j] = params[kOffB1 + j];
  vec[kMlpHid + j] = params[kOffB2 + j];
  vec[2 * kMlpHid + j] = params[kOffW3 + j];
  if (j == 0) vec[3 * kMlpHid] = params[kOffB3];
}

static hipError_t fold_into(const Mlp &m, const MlpInfer &dst, hipStream_t st) {
  hipLaunchKernelGGL(k_fold, dim3(kMlpIn), dim3(kMlpHid), 0, st, m.params, dst.ew1t);
  hipLaunchKernelGGL(k_inf_vec, dim3(1), dim3(kMlpHid), 0, st, m.params, dst.vec);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(dst.w2t, m.w2t, (uint64_t)kMlpHid * kMlpHid * 2, hipMemcpyDeviceToDevice,
                       st);
  return e;
}

hipError_t mlp_fold(Mlp &m, hipStream_t st) {
  hipError_t e = fold_into(m, m.live, st);
  if (e == hipSuccess) m.folded_valid = true;
  return e;
}

hipError_t mlp_fold_next(Mlp &m, hipStream_t st) { return fold_into(m, m.next, st); }

hipError_t mlp_publish(Mlp &m, hipStream_t st) {
  hipError_t e = hipMemcpyAsync(m.live.ew1t, m.next.ew1t, (uint64_t)kMlpIn * kMlpHid * 2,
                                hipMemcpyDeviceToDevice, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(m.live.w2t, m.next.w2t, (uint64_t)kMlpHid * kMlpHid * 2,
                       hipMemcpyDeviceToDevice, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(m.live.vec, m.next.vec, kInfVec * 4, hipMemcpyDeviceToDevice, st);
  return e;
}

static hipError_t refresh_bf16(Mlp &m, hipStream_t st) {
  m.folded_valid = false;
  hipLaunchKernelGGL(k_to_bf16, dim3(kMlpIn), dim3(kMlpHid), 0, st, m.params + kOffE, kMlpIn,
                     kMlpHid, m.e_w, m.e_wt);
  hipLaunchKernelGGL(k_to_bf16, dim3(kMlpHid), dim3(kMlpHid), 0, st, m.params + kOffW1,
                     kMlpHid, kMlpHid, m.w1, m.w1t);
  hipLaunchKernelGGL(k_to_bf16, dim3(kMlpHid), dim3(kMlpHid), 0, st, m.params + kOffW2,
                     kMlpHid, kMlpHid, m.w2, m.w2t);
  return hipGetLastError();
}

hipError_t mlp_init(Mlp &m, uint64_t seed, hipStream_t st) {
  std::vector<float> h(kNumParams, 0.f);
  uint64_t ctr = splitmix64(seed);
  auto uni = [&](float s) {
    const uint64_t r = splitmix64(ctr++);
    return s * (2.0f * (float)((r >> 11) * (1.0 / 9007199254740992.0)) - 1.0f);
  };
  const float se = std::sqrt(6.0f / (kMlpIn / 4 + kMlpHid));  // inputs are sparse histograms
  const float sh = std::sqrt(6.0f / (2 * kMlpHid));
  for (uint64_t i = 0; i < (uint64_t)kMlpIn * kMlpHid; i++) h[kOffE + i] = uni(se);
  for (uint64_t i = 0; i < (uint64_t)kMlpHid * kMlpHid; i++) h[kOffW1 + i] = uni(sh);
  for (uint64_t i = 0; i < (uint64_t)kMlpHid * kMlpHid; i++) h[kOffW2 + i] = uni(sh);
  for (int i = 0; i < kMlpHid; i++) h[kOffW3 + i] = uni(std::sqrt(6.0f / (kMlpHid + 1)));
  h[kOffB3] = 4.0f;  // typical log2(ops + 1)
  return mlp_set_params(m, h.data(), st);
}

hipError_t mlp_set_params(Mlp &m, const float *host, hipStream_t st) {
  hipError_t e = hipMemcpyAsync(m.params, host, kNumParams * 4, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemsetAsync(m.m, 0, kNumParams * 4, st);
  if (e == hipSuccess) e = hipMemsetAsync(m.v, 0, kNumParams * 4, st);
  m.adam_t = 0;
  if (e == hipSuccess) e = refresh_bf16(m, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  return e;
}

hipError_t mlp_get_params(const Mlp &m, float *host, hipStream_t st) {
  hipError_t e = hipMemcpyAsync(host, m.params, kNumParams * 4, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  return e;
}

hipError_t mlp_features(Mlp &m, const uint8_t *tapes, uint64_t count, hipStream_t st) {
  return mlp_features_rows(m, tapes, nullptr, count, st);
}

hipError_t mlp_features_rows(Mlp &m, const uint8_t *stage, const uint32_t *rows,
                             uint64_t count, hipStream_t st) {
  if (count > m.capacity) return hipErrorInvalidValue;
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_features, dim3((unsigned)((count + 15) / 16)), dim3(256), 0, st, stage,
                     rows, count, m.x);
  return hipGetLastError();
}

hipError_t mlp_forward(Mlp &m, uint64_t count, hipStream_t st) {
  if (count > m.capacity) return hipErrorInvalidValue;
  const uint64_t M = pad128(count);
  if (M == 0) return hipSuccess;
  const dim3 grid((unsigned)(M / 128)), block(512);
  // p = x E;  a1 = relu(p W1 + b1);  a2 = relu(a1 W2 + b2)
  hipLaunchKernelGGL((k_gemm_fwd<kMlpIn, kEpiBf16>), grid, block, 0, st, m.x, m.e_wt,
                     nullptr, m.p);
  hipLaunchKernelGGL((k_gemm_fwd<kMlpHid, kEpiBiasRelu>), grid, block, 0, st, m.p, m.w1t,
                     m.params + kOffB1, m.a1);
  hipLaunchKernelGGL((k_gemm_fwd<kMlpHid, kEpiBiasRelu>), grid, block, 0, st, m.a1, m.w2t,
                     m.params + kOffB2, m.a2);
  hipLaunchKernelGGL(k_out, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, m.a2,
                     m.params, m.y, count);
  return hipGetLastError();
}

hipError_t mlp_backward(Mlp &m, uint64_t count, hipStream_t st) {
  if (count > m.capacity || count == 0) return hipErrorInvalidValue;
  const uint64_t M = pad64(count);
  hipError_t e;
  if ((e = hipMemsetAsync(m.loss, 0, 4, st))) return e;
  if ((e = hipMemsetAsync(m.grads, 0, kNumParams * 4, st))) return e;  // all accumulated
  const int sk = wgrad_splits(M);
  const dim3 red((unsigned)((count + kRedRows - 1) / kRedRows));
  hipLaunchKernelGGL(k_out_bwd, dim3((unsigned)M), dim3(kMlpHid), 0, st, m.a2, m.params, m.y,
                     m.target, m.d2, count);
  hipLaunchKernelGGL(k_out_wgrad, red, dim3(kMlpHid), 0, st, m.a2, m.y, count, m.grads,
                     m.loss);
  // dW2 = a1^T d2, db2
  if ((e = gemm<false, false, kEpiF32Add>(st, kMlpHid, kMlpHid, (int)M, m.a1, 1, kMlpHid,
                                          m.d2, kMlpHid, 1, nullptr, m.grads + kOffW2, kMlpHid,
                                          nullptr, sk)))
    return e;
  hipLaunchKernelGGL(k_bias_grad, red, dim3(kMlpHid), 0, st, m.d2, count, m.grads + kOffB2);
  // d1 = (d2 W2^T) * relu'(a1)
  if ((e = gemm<true, true, kEpiMaskBf16>(st, (int)M, kMlpHid, kMlpHid, m.d2, kMlpHid, 1, m.w2,
                                          1, kMlpHid, nullptr, m.d1, kMlpHid, m.a1)))
    return e;
  // dW1 = p^T d1, db1
  if ((e = gemm<false, false, kEpiF32Add>(st, kMlpHid, kMlpHid, (int)M, m.p, 1, kMlpHid,
                                          m.d1, kMlpHid, 1, nullptr, m.grads + kOffW1, kMlpHid,
                                          nullptr, sk)))
    return e;
  hipLaunchKernelGGL(k_bias_grad, red, dim3(kMlpHid), 0, st, m.d1, count, m.grads + kOffB1);
  // dp = d1 W1^T
  if ((e = gemm<true, true, kEpiBf16>(st, (int)M, kMlpHid, kMlpHid, m.d1, kMlpHid, 1, m.w1, 1,
                                      kMlpHid, nullptr, m.dp, kMlpHid, nullptr)))
    return e;
  // dE = x^T dp
  if ((e = gemm<false, false, kEpiF32Add>(st, kMlpIn, kMlpHid, (int)M, m.x, 1, kMlpIn, m.dp,
                                          kMlpHid, 1, nullptr, m.grads + kOffE, kMlpHid,
                                          nullptr, sk)))
    return e;
  return hipGetLastError();
}

hipError_t mlp_adam(Mlp &m, float lr, hipStream_t st) {
  m.adam_t++;
  const float c1 = 1.0f - std::pow(0.9f, (float)m.adam_t);
  const float c2 = 1.0f - std::pow(0.999f, (float)m.adam_t);
  hipLaunchKernelGGL(k_adam, dim3((unsigned)((kNumParams + 255) / 256)), dim3(256), 0, st,
                     m.params, m.grads, m.m, m.v, kNumParams, lr, c1, c2);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = refresh_bf16(m, st);
  return e;
}

hipError_t mlp_targets_from_ops(Mlp &m, const uint32_t *ops, const uint32_t *rows,
                                uint64_t count, hipStream_t st) {
  if (count > m.capacity) return hipErrorInvalidValue;
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_targets, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, ops,
                     rows, count, m.target);
  return hipGetLastError();
}

}  // namespace cubff
