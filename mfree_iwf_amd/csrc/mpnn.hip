// Message-passing surrogate: forward and backward kernels on gfx950.
//
// Dense layers: k_gemm computes out[n][m] = sum_k X[n][k] W[m][k] with
// v_mfma_f32_16x16x32_bf16.  Operands are 8 contiguous bf16 per lane (A row
// = lane%16, B column = lane%16; result column = lane%16, rows
// 4*(lane/16)+v).  A workgroup stages the whole weight (at most 128 x 224
// bf16, rows padded by 16 bytes against LDS bank conflicts) in LDS once and
// then walks node tiles (persistent grid); a wave owns ROWS x 16 nodes and
// every 16-wide output tile, so each weight fragment read from LDS feeds
// ROWS MFMAs.  Epilogues fuse bias, SiLU, residual and the FP32 / BF16
// stores.  Weight gradients (k_gemm_dw) use the same instruction with the
// node axis as the reduction axis and per-chunk partial sums.  The
// aggregation over the neighbour graph is a gather per node (one node per
// lane, all 32 channels in registers) that evaluates the edge weights from
// the node positions; its backward is the same gather with the signs of the
// antisymmetric gradient weights flipped.
#include "mpnn.h"

#include <math.h>
#include <string.h>

#include <algorithm>
#include <random>

namespace mfree {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline float sigm(float z) { return 1.f / (1.f + __expf(-z)); }
__device__ inline float silu(float z) { return z * sigm(z); }
__device__ inline float silu_d(float z) {
  const float s = sigm(z);
  return s * (1.f + z * (1.f - s));
}

enum Epi : int {
  EPI_F32 = 0,       // out = acc + bias
  EPI_F32_ACC = 1,   // out += acc
  EPI_BF16 = 2,      // out16 = bf16(acc + bias)
  EPI_SILU_RES = 3,  // z = acc + bias -> zout; h = res + silu(z) -> out (f32), out16 (bf16)
  EPI_SPLIT = 4,     // columns < kH: out[n][m] += acc (ld kH); others: out16[n][m - kH] = bf16(acc) (ld 3 kC)
};

struct GemmEpi {
  int mode;
  const float* bias;            // nbias entries (columns beyond get no bias)
  int nbias;
  float* out;                   // FP32 output (or null)
  int ldo;
  bf16_t* out16;                // BF16 output
  int ld16;
  float* zout;                  // pre-activation (EPI_SILU_RES, or null)
  const float* res;             // FP32 residual (ld kH), or
  const bf16_t* res16;          // BF16 residual (ld kG), or neither
};

constexpr int kGemmBlocksMax = 512;  // 2 per CU
constexpr int kGemmThreads = 512;    // 8 waves share the LDS weight tile

__device__ inline uint2 pack4_bf16(float a, float b, float c, float d) {
  return make_uint2((uint32_t)f2bf(a) | ((uint32_t)f2bf(b) << 16), (uint32_t)f2bf(c) | ((uint32_t)f2bf(d) << 16));
}
__device__ inline float4 unpack4_bf16(uint2 u) {
  return make_float4(bf2f((bf16_t)(u.x & 0xFFFF)), bf2f((bf16_t)(u.x >> 16)), bf2f((bf16_t)(u.y & 0xFFFF)),
                     bf2f((bf16_t)(u.y >> 16)));
}

// out = X W^T over N rows; W is M x K (bf16, row-major, K % 32 == 0,
// M % 16 == 0).  The MFMA computes the transposed tile D = W X^T: A = 16
// weight rows, B = 16 node rows (both 8 contiguous bf16 per lane), so a lane
// ends with one node and 4 consecutive output channels (vector epilogue).
// Persistent: block b handles node tiles b, b + grid, ...
template <int K, int M, int ROWS, int MODE>
__global__ __launch_bounds__(kGemmThreads) void k_gemm(const bf16_t* __restrict__ X, int ldx,
                                                       const bf16_t* __restrict__ W, uint32_t N, GemmEpi e) {
  constexpr int KP = K + 8;  // padded LDS row
  constexpr int T = M / 16;
  constexpr int kWaves = kGemmThreads / 64;
  constexpr uint32_t kRowsPerBlock = kWaves * ROWS * 16;
  __shared__ __attribute__((aligned(16))) bf16_t wl[M * KP];
  for (int i = threadIdx.x * 8; i < M * K; i += kGemmThreads * 8) {
    const int r = i / K, c = i - r * K;
    *(uint4*)&wl[r * KP + c] = *(const uint4*)&W[i];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int kq = (lane >> 4) * 8;
  const int mq = (lane >> 4) * 4;  // first of the lane's 4 output channels within a 16-wide tile
  for (uint32_t tile = blockIdx.x; tile * kRowsPerBlock < N; tile += gridDim.x) {
    const uint32_t nb = tile * kRowsPerBlock + wave * ROWS * 16;
    if (nb >= N) continue;
    f32x4 acc[ROWS][T];
#pragma unroll
    for (int r = 0; r < ROWS; r++)
#pragma unroll
      for (int t = 0; t < T; t++) acc[r][t] = f32x4{0.f, 0.f, 0.f, 0.f};
    // all node-row fragments of the tile in flight before the first MFMA
    bf16x8 xb[K / 32][ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; r++) {
      const uint32_t row = nb + r * 16 + (lane & 15);
#pragma unroll
      for (int kk = 0; kk < K / 32; kk++)
        xb[kk][r] = row < N ? *(const bf16x8*)(X + (size_t)row * ldx + kk * 32 + kq) : bf16x8{};
    }
#pragma unroll
    for (int kk = 0; kk < K / 32; kk++) {
#pragma unroll
      for (int t = 0; t < T; t++) {
        const bf16x8 wa = *(const bf16x8*)&wl[(t * 16 + (lane & 15)) * KP + kk * 32 + kq];
#pragma unroll
        for (int r = 0; r < ROWS; r++)
          acc[r][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa, xb[kk][r], acc[r][t], 0, 0, 0);
      }
    }
#pragma unroll
    for (int r = 0; r < ROWS; r++) {
      const uint32_t n = nb + r * 16 + (lane & 15);
      if (n >= N) continue;
#pragma unroll
      for (int t = 0; t < T; t++) {
        const int m = t * 16 + mq;
        float4 x = make_float4(acc[r][t][0], acc[r][t][1], acc[r][t][2], acc[r][t][3]);
        if (e.bias) {
          x.x += m < e.nbias ? e.bias[m] : 0.f;
          x.y += m + 1 < e.nbias ? e.bias[m + 1] : 0.f;
          x.z += m + 2 < e.nbias ? e.bias[m + 2] : 0.f;
          x.w += m + 3 < e.nbias ? e.bias[m + 3] : 0.f;
        }
        if constexpr (MODE == EPI_F32) {
          *(float4*)&e.out[(size_t)n * e.ldo + m] = x;
        } else if constexpr (MODE == EPI_F32_ACC) {
          float4* o = (float4*)&e.out[(size_t)n * e.ldo + m];
          const float4 y = *o;
          *o = make_float4(y.x + x.x, y.y + x.y, y.z + x.z, y.w + x.w);
        } else if constexpr (MODE == EPI_BF16) {
          *(uint2*)&e.out16[(size_t)n * e.ld16 + m] = pack4_bf16(x.x, x.y, x.z, x.w);
        } else if constexpr (MODE == EPI_SPLIT) {
          if (m < kH) {
            float4* o = (float4*)&e.out[(size_t)n * kH + m];
            const float4 y = *o;
            *o = make_float4(y.x + x.x, y.y + x.y, y.z + x.z, y.w + x.w);
          } else {
            *(uint2*)&e.out16[(size_t)n * (3 * kC) + (m - kH)] = pack4_bf16(x.x, x.y, x.z, x.w);
          }
        } else {
          if (e.zout) *(float4*)&e.zout[(size_t)n * kH + m] = x;
          float4 h = make_float4(silu(x.x), silu(x.y), silu(x.z), silu(x.w));
          if (e.res) {
            const float4 rr = *(const float4*)&e.res[(size_t)n * kH + m];
            h = make_float4(h.x + rr.x, h.y + rr.y, h.z + rr.z, h.w + rr.w);
          } else if (e.res16) {
            const float4 rr = unpack4_bf16(*(const uint2*)&e.res16[(size_t)n * kG + m]);
            h = make_float4(h.x + rr.x, h.y + rr.y, h.z + rr.z, h.w + rr.w);
          }
          if (e.out) *(float4*)&e.out[(size_t)n * kH + m] = h;
          *(uint2*)&e.out16[(size_t)n * e.ld16 + m] = pack4_bf16(h.x, h.y, h.z, h.w);
        }
      }
    }
  }
}

// Weight gradient partials: part[chunk][m][k] = sum_{n in chunk} dY[n][m] X[n][k]
// for an M x K weight; grid (M/16, chunks), 4 waves per block split the
// chunk's nodes, the block reduces them through LDS.
constexpr uint32_t kDwChunk = 8192;
template <int K>
__global__ __launch_bounds__(256) void k_gemm_dw(const bf16_t* __restrict__ dY, int ldy, const bf16_t* __restrict__ X,
                                                 int ldx, uint32_t N, int M, float* __restrict__ part) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m0 = blockIdx.x * 16;
  const uint32_t c0 = blockIdx.y * kDwChunk;
  const uint32_t c1 = min(N, c0 + kDwChunk);
  constexpr int T = K / 16;
  f32x4 acc[T];
#pragma unroll
  for (int t = 0; t < T; t++) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int nq = (lane >> 4) * 8;
  for (uint32_t n0 = c0 + wave * 32; n0 < c1; n0 += 128) {
    bf16x8 a;
#pragma unroll
    for (int e = 0; e < 8; e++) {
      const uint32_t n = n0 + nq + e;
      bf16_t v = 0;
      if (n < c1) v = dY[(size_t)n * ldy + m0 + (lane & 15)];
      a[e] = __builtin_bit_cast(__bf16, v);
    }
#pragma unroll
    for (int t = 0; t < T; t++) {
      bf16x8 b;
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const uint32_t n = n0 + nq + e;
        bf16_t v = 0;
        if (n < c1) v = X[(size_t)n * ldx + t * 16 + (lane & 15)];
        b[e] = __builtin_bit_cast(__bf16, v);
      }
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
    }
  }
  // reduce the 4 waves: result tile [16 x K]
  __shared__ float red[4][16 * 64];
#pragma unroll
  for (int tb = 0; tb < T; tb += 4) {
    const int tn = min(4, T - tb);
#pragma unroll
    for (int tt = 0; tt < 4; tt++) {
      if (tb + tt < T) {
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int r = (lane >> 4) * 4 + v, cc = tt * 16 + (lane & 15);
          red[wave][r * 64 + cc] = acc[tb + tt][v];
        }
      }
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < 16 * tn * 16; idx += 256) {
      const int r = idx / (tn * 16), cc = idx % (tn * 16);
      const float s = red[0][r * 64 + cc] + red[1][r * 64 + cc] + red[2][r * 64 + cc] + red[3][r * 64 + cc];
      part[((size_t)blockIdx.y * M + m0 + r) * K + tb * 16 + cc] = s;
    }
    __syncthreads();
  }
}

// g[m][k] += sum over chunks of part[chunk][m][k]
__global__ void k_reduce_parts(const float* __restrict__ part, uint32_t nchunk, int MK, float* __restrict__ g) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= MK) return;
  float s = 0.f;
  for (uint32_t c = 0; c < nchunk; c++) s += part[(size_t)c * MK + i];
  g[i] += s;
}

// g[r][k] += sum_chunks part[chunk][r][k] for r < rows, k < cols of a
// partial laid out [chunk][Mp][Kp] (the padded GEMM shape).
__global__ void k_reduce_rect(const float* __restrict__ part, uint32_t nchunk, int Mp, int Kp, int rows, int cols,
                              float* __restrict__ g) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * cols) return;
  const int r = i / cols, k = i % cols;
  float s = 0.f;
  for (uint32_t c = 0; c < nchunk; c++) s += part[((size_t)c * Mp + r) * Kp + k];
  g[i] += s;
}

// bias gradient: g[m] += sum_n dY[n][m] (bf16 operand, M <= 256 columns)
__global__ __launch_bounds__(256) void k_colsum(const bf16_t* __restrict__ dY, int ldy, uint32_t N, int M,
                                                float* __restrict__ part) {
  const uint32_t c0 = blockIdx.x * kDwChunk, c1 = min(N, c0 + kDwChunk);
  const int m = threadIdx.x;
  if (m >= M) return;
  float s = 0.f;
  for (uint32_t n = c0; n < c1; n++) s += bf2f(dY[(size_t)n * ldy + m]);
  part[(size_t)blockIdx.x * M + m] = s;
}

__device__ inline void unpack8_bf16(const uint4 u, float* f) {
  const uint32_t w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
  for (int k = 0; k < 4; k++) {
    f[2 * k] = bf2f((bf16_t)(w[k] & 0xFFFF));
    f[2 * k + 1] = bf2f((bf16_t)(w[k] >> 16));
  }
}
__device__ inline uint4 pack8_bf16(const float* f) {
  uint32_t w[4];
#pragma unroll
  for (int k = 0; k < 4; k++) w[k] = (uint32_t)f2bf(f[2 * k]) | ((uint32_t)f2bf(f[2 * k + 1]) << 16);
  return make_uint4(w[0], w[1], w[2], w[3]);
}

// Aggregation forward, one node per lane (consecutive nodes on consecutive
// lanes, so the slot-major edge lists load coalesced):
//   G[i][128+c] = sum_e w U[j][c], G[i][160+c] = sum_e gx U[j][c], G[i][192+c] = sum_e gy U[j][c]
__global__ __launch_bounds__(256) void k_agg_fwd(uint32_t n, uint32_t stride, const uint32_t* __restrict__ deg,
                                                 const uint32_t* __restrict__ col, const float4* __restrict__ npos,
                                                 const bf16_t* __restrict__ U, bf16_t* __restrict__ G) {
  const uint32_t i = xcd_tile() * 256u + threadIdx.x;
  if (i >= n) return;
  const float4 pi = npos[i];
  float a0[kC], ax[kC], ay[kC];
#pragma unroll
  for (int c = 0; c < kC; c++) a0[c] = ax[c] = ay[c] = 0.f;
  const uint32_t d = deg[i];
  for (uint32_t k = 0; k < d; k++) {
    const uint32_t j = col[(size_t)k * stride + i];
    const float4 pj = npos[j];
    const uint4* u = (const uint4*)(U + (size_t)j * kC);
    uint4 uv[kC / 8];
#pragma unroll
    for (int q = 0; q < kC / 8; q++) uv[q] = u[q];
    float w, gx, gy;
    edge_weights(pi, pj, w, gx, gy);
#pragma unroll
    for (int q = 0; q < kC / 8; q++) {
      float f[8];
      unpack8_bf16(uv[q], f);
#pragma unroll
      for (int c = 0; c < 8; c++) {
        a0[8 * q + c] += w * f[c];
        ax[8 * q + c] += gx * f[c];
        ay[8 * q + c] += gy * f[c];
      }
    }
  }
  uint4* g = (uint4*)(G + (size_t)i * kG + kH);
#pragma unroll
  for (int q = 0; q < kC / 8; q++) {
    g[q] = pack8_bf16(a0 + 8 * q);
    g[kC / 8 + q] = pack8_bf16(ax + 8 * q);
    g[2 * (kC / 8) + q] = pack8_bf16(ay + 8 * q);
  }
}

// Aggregation backward, one node per lane:
//   dU[j][c] = sum_{i in N(j)} (w dA0_i - gx_ji dAx_i - gy_ji dAy_i)
// (w symmetric, gx/gy antisymmetric); dA = dA16 [n x 96] bf16.  Writes dU
// as bf16 into dU16 (ld ldu).
__global__ __launch_bounds__(256) void k_agg_bwd(uint32_t n, uint32_t stride, const uint32_t* __restrict__ deg,
                                                 const uint32_t* __restrict__ col, const float4* __restrict__ npos,
                                                 const bf16_t* __restrict__ dA16, bf16_t* __restrict__ dU16, int ldu) {
  const uint32_t j = xcd_tile() * 256u + threadIdx.x;
  if (j >= n) return;
  const float4 pj = npos[j];
  float s[kC];
#pragma unroll
  for (int c = 0; c < kC; c++) s[c] = 0.f;
  const uint32_t d = deg[j];
  for (uint32_t k = 0; k < d; k++) {
    const uint32_t i = col[(size_t)k * stride + j];
    const float4 pi = npos[i];
    float w, gx, gy;
    edge_weights(pj, pi, w, gx, gy);  // (gx, gy)_ji = -(gx, gy)_ij
    const uint4* da = (const uint4*)(dA16 + (size_t)i * (3 * kC));
#pragma unroll
    for (int q = 0; q < kC / 8; q++) {
      float f0[8], f1[8], f2[8];
      unpack8_bf16(da[q], f0);
      unpack8_bf16(da[kC / 8 + q], f1);
      unpack8_bf16(da[2 * (kC / 8) + q], f2);
#pragma unroll
      for (int c = 0; c < 8; c++) s[8 * q + c] += w * f0[c] - gx * f1[c] - gy * f2[c];
    }
  }
  uint4* o = (uint4*)(dU16 + (size_t)j * ldu);
#pragma unroll
  for (int q = 0; q < kC / 8; q++) o[q] = pack8_bf16(s + 8 * q);
}

// dZ = dH * silu'(Z) -> bf16 operand (ld 224 buffer, columns [0,128))
__global__ void k_silu_bwd(uint32_t nelem, const float* __restrict__ dH, const float* __restrict__ Z,
                           bf16_t* __restrict__ dZ16, int ld16) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nelem) return;
  const uint32_t n = idx / kH, m = idx % kH;
  dZ16[(size_t)n * ld16 + m] = f2bf(dH[idx] * silu_d(Z[idx]));
}

// MSE loss over the kOut outputs: dY = 2 (Y - T) / (kOut N) as bf16 (ld 32,
// columns kOut..31 zero); loss accumulated into *loss (device float).
__global__ void k_mse(uint32_t N, const float* __restrict__ Y, const float* __restrict__ T, bf16_t* __restrict__ dY16,
                      float* __restrict__ loss) {
  const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (n < N) {
    const float inv = 2.f / ((float)kOut * (float)N);
    for (int k = 0; k < 32; k++) {
      float d = 0.f;
      if (k < kOut) {
        const float r = Y[(size_t)n * kOutP + k] - T[(size_t)n * kOut + k];
        l += r * r;
        d = inv * r;
      }
      dY16[(size_t)n * 32 + k] = f2bf(d);
    }
  }
  for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd(loss, l / ((float)kOut * (float)N));
}

__global__ void k_fill_f32(float* p, float v, size_t n) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Adam (Kingma & Ba 2015) on the flat FP32 master weights.
__global__ void k_adam(size_t n, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, float lr, float b1, float b2, float eps, float c1, float c2) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.f - b1) * gi;
  const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  w[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// BF16 copies of a weight: W [rows x cols] and W^T [cols x rows_p] (rows
// zero-padded to rows_p).
__global__ void k_sync16(const float* __restrict__ w, int rows, int cols, int cols_p, bf16_t* __restrict__ w16,
                         bf16_t* __restrict__ wt16, int rows_p) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= rows * cols_p) return;
  const int r = idx / cols_p, c = idx % cols_p;
  const float v = c < cols ? w[(size_t)r * cols + c] : 0.f;
  w16[(size_t)r * cols_p + c] = f2bf(v);
  if (wt16 && c < cols) wt16[(size_t)c * rows_p + r] = f2bf(v);
}

inline uint32_t blocks(size_t n, int b) { return (uint32_t)((n + b - 1) / b); }

// Rows per wave: accumulators ROWS x M/16 x 4 registers stay within budget.
template <int K, int M, int ROWS = (M > 128 ? 1 : (M > 32 ? 2 : 4))>
void gemm(const bf16_t* X, int ldx, const bf16_t* W, uint32_t N, const GemmEpi& e, hipStream_t st) {
  const uint32_t tiles = blocks(N, (kGemmThreads / 64) * ROWS * 16);
  const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(tiles, kGemmBlocksMax));
  switch (e.mode) {
    case EPI_F32:
      hipLaunchKernelGGL((k_gemm<K, M, ROWS, EPI_F32>), dim3(grid), dim3(kGemmThreads), 0, st, X, ldx, W, N, e);
      break;
    case EPI_F32_ACC:
      hipLaunchKernelGGL((k_gemm<K, M, ROWS, EPI_F32_ACC>), dim3(grid), dim3(kGemmThreads), 0, st, X, ldx, W, N, e);
      break;
    case EPI_BF16:
      hipLaunchKernelGGL((k_gemm<K, M, ROWS, EPI_BF16>), dim3(grid), dim3(kGemmThreads), 0, st, X, ldx, W, N, e);
      break;
    case EPI_SPLIT:
      hipLaunchKernelGGL((k_gemm<K, M, ROWS, EPI_SPLIT>), dim3(grid), dim3(kGemmThreads), 0, st, X, ldx, W, N, e);
      break;
    default:
      hipLaunchKernelGGL((k_gemm<K, M, ROWS, EPI_SILU_RES>), dim3(grid), dim3(kGemmThreads), 0, st, X, ldx, W, N, e);
      break;
  }
}

GemmEpi epi(int mode) {
  GemmEpi e;
  memset(&e, 0, sizeof(e));
  e.mode = mode;
  return e;
}

// g[M x K] += dY^T X over all N nodes
template <int K>
void gemm_dw(const bf16_t* dY, int ldy, const bf16_t* X, int ldx, uint32_t N, int M, float* part, float* g,
             hipStream_t st) {
  const uint32_t nch = blocks(N, kDwChunk);
  hipLaunchKernelGGL((k_gemm_dw<K>), dim3(M / 16, nch), dim3(256), 0, st, dY, ldy, X, ldx, N, M, part);
  hipLaunchKernelGGL(k_reduce_parts, dim3(blocks((size_t)M * K, 256)), dim3(256), 0, st, part, nch, M * K, g);
}

void bias_grad(const bf16_t* dY, int ldy, uint32_t N, int M, float* part, float* g, hipStream_t st) {
  const uint32_t nch = blocks(N, kDwChunk);
  hipLaunchKernelGGL(k_colsum, dim3(nch), dim3(256), 0, st, dY, ldy, N, M, part);
  hipLaunchKernelGGL(k_reduce_parts, dim3(blocks(M, 256)), dim3(256), 0, st, part, nch, M, g);
}

}  // namespace

// ---------------------------------------------------------------------------

mfree_status mp_net_init(MpNet& net, uint64_t seed, std::string& msg) {
  net.t.clear();
  size_t off = 0;
  auto add = [&](int rows, int cols) {
    MpTensor t;
    t.rows = rows;
    t.cols = cols;
    t.off = off;
    off += (size_t)rows * cols;
    net.t.push_back(t);
    return (int)net.t.size() - 1;
  };
  net.We = add(kH, kNF);
  net.be = add(1, kH);
  for (int l = 0; l < kLayers; l++) {
    net.Wu[l] = add(kC, kH);
    net.Wo[l] = add(kH, kG);
    net.bo[l] = add(1, kH);
  }
  net.Wd = add(kOut, kH);
  net.bd = add(1, kOut);
  net.nparam = off;
  // He-style initialisation from a seeded generator; the last layer of every
  // residual branch is scaled down so the untrained net starts near identity
  std::vector<float> h(off, 0.f);
  std::mt19937_64 rng(seed);
  std::normal_distribution<float> nd(0.f, 1.f);
  for (size_t k = 0; k < net.t.size(); k++) {
    const MpTensor& t = net.t[k];
    if (t.rows == 1) continue;  // biases start at zero
    float sc = sqrtf(2.f / (float)t.cols);
    bool is_out = false;
    for (int l = 0; l < kLayers; l++) is_out |= (int)k == net.Wo[l];
    if (is_out) sc *= 0.3f;
    if ((int)k == net.Wd) sc *= 0.1f;
    for (size_t i = 0; i < (size_t)t.rows * t.cols; i++) h[t.off + i] = sc * nd(rng);
  }
  auto al = [&](void** p, size_t bytes) -> bool {
    if (hipMalloc(p, std::max<size_t>(bytes, 16)) != hipSuccess) return false;
    net.allocs.push_back(*p);
    return true;
  };
  if (!al((void**)&net.w, off * 4) || !al((void**)&net.g, off * 4) || !al((void**)&net.m1, off * 4) ||
      !al((void**)&net.m2, off * 4)) {
    msg = "hipMalloc failed";
    return MFREE_ERR_HIP;
  }
  for (auto& t : net.t) {
    if (t.rows == 1) continue;
    const int cols_p = (t.cols + 31) / 32 * 32;
    const int rows_p = (t.rows + 15) / 16 * 16;
    const int rows_pk = (t.rows + 31) / 32 * 32;  // W^T is the operand of a K-padded GEMM
    if (!al((void**)&t.w16, (size_t)rows_p * cols_p * 2) || !al((void**)&t.wt16, (size_t)t.cols * rows_pk * 2)) {
      msg = "hipMalloc failed";
      return MFREE_ERR_HIP;
    }
    (void)hipMemset(t.w16, 0, (size_t)rows_p * cols_p * 2);
    (void)hipMemset(t.wt16, 0, (size_t)t.cols * rows_pk * 2);
  }
  if (hipMemcpy(net.w, h.data(), off * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(net.m1, 0, off * 4) != hipSuccess || hipMemset(net.m2, 0, off * 4) != hipSuccess ||
      hipMemset(net.g, 0, off * 4) != hipSuccess) {
    msg = "hipMemcpy failed";
    return MFREE_ERR_HIP;
  }
  net.adam_t = 0;
  mp_net_sync16(net, 0);
  if (hipDeviceSynchronize() != hipSuccess) {
    msg = "init failed";
    return MFREE_ERR_HIP;
  }
  return MFREE_OK;
}

void mp_net_free(MpNet& net) {
  for (void* p : net.allocs) (void)hipFree(p);
  net.allocs.clear();
}

void mp_net_sync16(MpNet& net, hipStream_t st) {
  for (auto& t : net.t) {
    if (t.rows == 1) continue;
    const int cols_p = (t.cols + 31) / 32 * 32;
    const int rows_pk = (t.rows + 31) / 32 * 32;
    hipLaunchKernelGGL(k_sync16, dim3(blocks((size_t)t.rows * cols_p, 256)), dim3(256), 0, st, net.w + t.off, t.rows,
                       t.cols, cols_p, t.w16, t.wt16, rows_pk);
  }
}

mfree_status mp_acts_reserve(MpActs& a, uint32_t n, bool train, std::string& msg) {
  if (n <= a.cap && (a.train || !train)) return MFREE_OK;
  // grow to cover both the previous and the new request
  n = std::max(n, a.cap);
  train = train || a.train;
  mp_acts_free(a);
  auto al = [&](void** p, size_t bytes) -> bool {
    if (hipMalloc(p, std::max<size_t>(bytes, 16)) != hipSuccess) return false;
    a.allocs.push_back(*p);
    return true;
  };
  const size_t N = n;
  bool ok = al((void**)&a.X0, N * kK0 * 2) && al((void**)&a.Y, N * kOutP * 4);
  for (int l = 0; l <= kLayers && ok; l++) ok = al((void**)&a.G[l], N * kG * 2);
  for (int l = 0; l < kLayers && ok; l++) ok = al((void**)&a.U[l], N * kC * 2);
  if (train) {
    ok = ok && al((void**)&a.dH, N * kH * 4) && al((void**)&a.dZ16, N * kG * 2) &&
         al((void**)&a.dA16, N * 3 * kC * 2);
    for (int l = 0; l <= kLayers && ok; l++) ok = al((void**)&a.Z[l], N * kH * 4) && al((void**)&a.Hf[l], N * kH * 4);
    a.part_cap = (size_t)blocks(N, kDwChunk) * kH * kG;
    ok = ok && al((void**)&a.part, a.part_cap * 4);
  }
  if (!ok) {
    mp_acts_free(a);
    msg = "hipMalloc failed (activations)";
    return MFREE_ERR_HIP;
  }
  a.cap = n;
  a.train = train;
  return MFREE_OK;
}

void mp_acts_free(MpActs& a) {
  for (void* p : a.allocs) (void)hipFree(p);
  a.allocs.clear();
  a.cap = 0;
  a.train = false;
  for (int l = 0; l <= kLayers; l++) a.Z[l] = a.Hf[l] = nullptr;
  a.dH = nullptr;
  a.dZ16 = a.dA16 = nullptr;
  a.part = nullptr;
}

void mp_forward(const MpNet& net, MpActs& a, const MpGraph& g, uint32_t n, bool train, hipStream_t st) {
  // encoder: H0 = SiLU(X0 We^T + be)
  GemmEpi e = epi(EPI_SILU_RES);
  e.bias = net.w + net.t[net.be].off;
  e.nbias = kH;
  e.zout = train ? a.Z[0] : nullptr;
  e.out = train ? a.Hf[0] : nullptr;
  e.out16 = a.G[0];
  e.ld16 = kG;
  gemm<kK0, kH>(a.X0, kK0, net.t[net.We].w16, n, e, st);
  for (int l = 0; l < kLayers; l++) {
    // U = H Wu^T
    GemmEpi eu = epi(EPI_BF16);
    eu.out16 = a.U[l];
    eu.ld16 = kC;
    gemm<kH, kC>(a.G[l], kG, net.t[net.Wu[l]].w16, n, eu, st);
    // A = aggregate(U) into G[l][:, 128:224]
    hipLaunchKernelGGL(k_agg_fwd, dim3(xcd_grid(blocks(n, 256))), dim3(256), 0, st, n, g.cap, g.deg, g.col, g.npos,
                       a.U[l], a.G[l]);
    // H_{l+1} = H_l + SiLU([H, A] Wo^T + bo)
    GemmEpi eo = epi(EPI_SILU_RES);
    eo.bias = net.w + net.t[net.bo[l]].off;
    eo.nbias = kH;
    eo.zout = train ? a.Z[l + 1] : nullptr;
    eo.res = train ? a.Hf[l] : nullptr;
    eo.res16 = train ? nullptr : a.G[l];
    eo.out = train ? a.Hf[l + 1] : nullptr;
    eo.out16 = a.G[l + 1];
    eo.ld16 = kG;
    gemm<kG, kH>(a.G[l], kG, net.t[net.Wo[l]].w16, n, eo, st);
  }
  GemmEpi ed = epi(EPI_F32);
  ed.bias = net.w + net.t[net.bd].off;
  ed.nbias = kOut;
  ed.out = a.Y;
  ed.ldo = kOutP;
  gemm<kH, kOutP>(a.G[kLayers], kG, net.t[net.Wd].w16, n, ed, st);
}

mfree_status mp_graph_reserve(MpGraph& g, uint32_t n, std::string& msg) {
  if (n <= g.cap) return MFREE_OK;
  mp_graph_free(g);
  auto al = [&](void** p, size_t bytes) -> bool {
    if (hipMalloc(p, std::max<size_t>(bytes, 16)) != hipSuccess) return false;
    g.allocs.push_back(*p);
    return true;
  };
  const size_t N = n;
  if (!al((void**)&g.deg, N * 4) || !al((void**)&g.col, N * kMaxDeg * 4) || !al((void**)&g.npos, N * 16) ||
      !al((void**)&g.wsum, N * 4)) {
    mp_graph_free(g);
    msg = "hipMalloc failed (graph)";
    return MFREE_ERR_HIP;
  }
  g.cap = n;
  return MFREE_OK;
}

void mp_graph_free(MpGraph& g) {
  for (void* p : g.allocs) (void)hipFree(p);
  g.allocs.clear();
  g.cap = 0;
}

void mp_zero_grad(MpNet& net, hipStream_t st) {
  hipLaunchKernelGGL(k_fill_f32, dim3(blocks(net.nparam, 256)), dim3(256), 0, st, net.g, 0.f, net.nparam);
}

void mp_backward(MpNet& net, MpActs& a, const MpGraph& g, uint32_t n, const float* T, float* loss_out, hipStream_t st) {
  // loss and dY (bf16, ld 32, zero-padded)
  hipLaunchKernelGGL(k_mse, dim3(blocks(n, 256)), dim3(256), 0, st, n, a.Y, T, a.dZ16, loss_out);
  // decoder: dWd = dY^T H3 (rows 0..kOut-1 of the padded 16), dbd = sum dY, dH3 = dY Wd
  const uint32_t nch = blocks(n, kDwChunk);
  hipLaunchKernelGGL((k_gemm_dw<kH>), dim3(1, nch), dim3(256), 0, st, a.dZ16, 32, a.G[kLayers], kG, n, 16, a.part);
  hipLaunchKernelGGL(k_reduce_rect, dim3(blocks((size_t)kOut * kH, 256)), dim3(256), 0, st, a.part, nch, 16, kH, kOut,
                     kH, net.g + net.t[net.Wd].off);
  bias_grad(a.dZ16, 32, n, kOut, a.part, net.g + net.t[net.bd].off, st);
  GemmEpi e = epi(EPI_F32);
  e.out = a.dH;
  e.ldo = kH;
  gemm<32, kH>(a.dZ16, 32, net.t[net.Wd].wt16, n, e, st);
  for (int l = kLayers - 1; l >= 0; l--) {
    // dZ = dH * silu'(Z_{l+1}); the residual keeps dH for H_l
    hipLaunchKernelGGL(k_silu_bwd, dim3(blocks((size_t)n * kH, 256)), dim3(256), 0, st, n * kH, a.dH, a.Z[l + 1],
                       a.dZ16, kG);
    // dWo = dZ^T G_l, dbo = sum dZ
    gemm_dw<kG>(a.dZ16, kG, a.G[l], kG, n, kH, a.part, net.g + net.t[net.Wo[l]].off, st);
    bias_grad(a.dZ16, kG, n, kH, a.part, net.g + net.t[net.bo[l]].off, st);
    // [dH_l += dG[:, :128], dA = dG[:, 128:224]] with dG = dZ Wo
    GemmEpi eg = epi(EPI_SPLIT);
    eg.out = a.dH;
    eg.out16 = a.dA16;
    gemm<kH, kG>(a.dZ16, kG, net.t[net.Wo[l]].wt16, n, eg, st);
    // dU = aggregate^T(dA) -> bf16 operand (dZ16 columns 0..31)
    hipLaunchKernelGGL(k_agg_bwd, dim3(xcd_grid(blocks(n, 256))), dim3(256), 0, st, n, g.cap, g.deg, g.col, g.npos,
                       a.dA16, a.dZ16, kG);
    // dWu = dU^T H_l, dH_l += dU Wu
    gemm_dw<kH>(a.dZ16, kG, a.G[l], kG, n, kC, a.part, net.g + net.t[net.Wu[l]].off, st);
    GemmEpi eu = epi(EPI_F32_ACC);
    eu.out = a.dH;
    eu.ldo = kH;
    gemm<kC, kH>(a.dZ16, kG, net.t[net.Wu[l]].wt16, n, eu, st);
  }
  // encoder: dZ0 = dH0 * silu'(Z0); dWe = dZ0^T X0 (first kNF columns), dbe
  hipLaunchKernelGGL(k_silu_bwd, dim3(blocks((size_t)n * kH, 256)), dim3(256), 0, st, n * kH, a.dH, a.Z[0], a.dZ16,
                     kG);
  hipLaunchKernelGGL((k_gemm_dw<kK0>), dim3(kH / 16, nch), dim3(256), 0, st, a.dZ16, kG, a.X0, kK0, n, kH, a.part);
  hipLaunchKernelGGL(k_reduce_rect, dim3(blocks((size_t)kH * kNF, 256)), dim3(256), 0, st, a.part, nch, kH, kK0, kH,
                     kNF, net.g + net.t[net.We].off);
  bias_grad(a.dZ16, kG, n, kH, a.part, net.g + net.t[net.be].off, st);
}

void mp_adam(MpNet& net, float lr, float beta1, float beta2, float eps, hipStream_t st) {
  net.adam_t++;
  const float c1 = 1.f - powf(beta1, (float)net.adam_t), c2 = 1.f - powf(beta2, (float)net.adam_t);
  hipLaunchKernelGGL(k_adam, dim3(blocks(net.nparam, 256)), dim3(256), 0, st, net.nparam, net.w, net.g, net.m1,
                     net.m2, lr, beta1, beta2, eps, c1, c2);
  mp_net_sync16(net, st);
}

}  // namespace mfree
