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
  const bf16_t* res16;          // BF16 residual (ld kH), or neither
  // EPI_SILU_RES with M == kH: also U = bf16(h) Wu^T of the next layer
  const bf16_t* wu;             // kC x kH bf16 (or null)
  bf16_t* u_out;                // N x kC bf16
};

// Launch shape: a weight tile above 32 KB is shared by 16 waves of one
// workgroup per CU; smaller tiles use 4-wave workgroups, several per CU.
template <int K, int M>
constexpr int gemm_threads() {
  return M * (K + 8) * 2 > 32 * 1024 ? 1024 : 256;
}
template <int K, int M>
constexpr int gemm_blocks_max() {
  return M * (K + 8) * 2 > 32 * 1024 ? 256 : 2048;
}

__device__ inline float4 unpack4_bf16(uint2 u) {
  return make_float4(bf2f((bf16_t)(u.x & 0xFFFF)), bf2f((bf16_t)(u.x >> 16)), bf2f((bf16_t)(u.y & 0xFFFF)),
                     bf2f((bf16_t)(u.y >> 16)));
}

// out = [X1, X2] W^T over N rows; W is M x K (bf16, row-major, K % 32 == 0,
// M % 16 == 0); columns k < KS of the operand come from X1 (ld ldx), the
// rest from X2 (ld ldx2).  The MFMA computes the transposed tile D = W X^T:
// A = 16 weight rows, B = 16 node rows (both 8 contiguous bf16 per lane),
// so a lane ends with one node and 4 consecutive output channels (vector
// epilogue).  Persistent: block b handles node tiles b, b + grid, ...
template <int K, int M, int ROWS, int MODE, int KS, int THREADS, bool FU = false>
__global__ __launch_bounds__(THREADS) void k_gemm(const bf16_t* __restrict__ X, int ldx,
                                                       const bf16_t* __restrict__ X2, int ldx2,
                                                       const bf16_t* __restrict__ W, uint32_t N, GemmEpi e) {
  constexpr int KP = K + 8;  // padded LDS row
  constexpr int T = M / 16;
  constexpr int kWaves = THREADS / 64;
  constexpr uint32_t kRowsPerBlock = kWaves * ROWS * 16;
  __shared__ __attribute__((aligned(16))) bf16_t wl[M * KP];
  // bf16 outputs leave through a per-wave LDS tile (16 rows x M) so that the
  // stores are whole rows of 16-byte chunks
  constexpr bool kStage = (MODE == EPI_BF16 || MODE == EPI_SILU_RES) && M <= 128;
  constexpr int MP = M + 8;
  __shared__ __attribute__((aligned(16))) bf16_t ostage[kStage ? kWaves * 16 * MP : 8];
  // fused message projection of the next layer (FU): Wu staged like W
  constexpr int KU = kH + 8;
  __shared__ __attribute__((aligned(16))) bf16_t wul[FU ? kC * KU : 8];
  if constexpr (FU) {
    for (int i = threadIdx.x * 8; i < kC * kH; i += THREADS * 8) {
      const int r = i / kH, c = i - r * kH;
      *(uint4*)&wul[r * KU + c] = *(const uint4*)&e.wu[i];
    }
  }
  for (int i = threadIdx.x * 8; i < M * K; i += THREADS * 8) {
    const int r = i / K, c = i - r * K;
    *(uint4*)&wl[r * KP + c] = *(const uint4*)&W[i];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int kq = (lane >> 4) * 8;
  const int mq = (lane >> 4) * 4;  // first of the lane's 4 output channels within a 16-wide tile
  // Operand rows of one tile; the next tile's are loaded before this tile's
  // epilogue stores, so waiting for them does not wait for the stores.
  bf16x8 xb[K / 32][ROWS];
#define MP_LOAD_TILE(TILE)                                                                                   \
  do {                                                                                                       \
    const uint32_t nb_ = (TILE) * kRowsPerBlock + wave * ROWS * 16;                                          \
    _Pragma("unroll") for (int r = 0; r < ROWS; r++) {                                                       \
      const uint32_t row = nb_ + r * 16 + (lane & 15);                                                       \
      _Pragma("unroll") for (int kk = 0; kk < K / 32; kk++) {                                                \
        const bf16_t* src = kk * 32 < KS ? X + (size_t)row * ldx + kk * 32 + kq                              \
                                         : X2 + (size_t)row * ldx2 + (kk * 32 - KS) + kq;                    \
        xb[kk][r] = row < N ? *(const bf16x8*)src : bf16x8{};                                                \
      }                                                                                                      \
    }                                                                                                        \
  } while (0)
  uint32_t tile = blockIdx.x;
  MP_LOAD_TILE(tile);
  for (; tile * kRowsPerBlock < N; tile += gridDim.x) {
    const uint32_t nb = tile * kRowsPerBlock + wave * ROWS * 16;
    f32x4 acc[ROWS][T];
#pragma unroll
    for (int r = 0; r < ROWS; r++)
#pragma unroll
      for (int t = 0; t < T; t++) acc[r][t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kk = 0; kk < K / 32; kk++) {
#pragma unroll
      for (int t = 0; t < T; t++) {
        const bf16x8 wa = *(const bf16x8*)&wl[(t * 16 + (lane & 15)) * KP + kk * 32 + kq];
#pragma unroll
        for (int r = 0; r < ROWS; r++)
          acc[r][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa, xb[kk][r], acc[r][t], 0, 0, 0);
      }
      // keep the weight fragments of one k-step live at a time
      __builtin_amdgcn_sched_barrier(0);
    }
    // residual rows of this tile, then the next tile's operands
    uint2 rr16[ROWS][T];
    if constexpr (MODE == EPI_SILU_RES) {
#pragma unroll
      for (int r = 0; r < ROWS; r++) {
        const uint32_t n = nb + r * 16 + (lane & 15);
#pragma unroll
        for (int t = 0; t < T; t++)
          rr16[r][t] = (e.res16 && n < N) ? *(const uint2*)&e.res16[(size_t)n * kH + t * 16 + mq] : make_uint2(0u, 0u);
      }
    }
    MP_LOAD_TILE(tile + gridDim.x);
#pragma unroll
    for (int r = 0; r < ROWS; r++) {
      const uint32_t n = nb + r * 16 + (lane & 15);
      bf16_t* stg = ostage + wave * 16 * MP;
      if (n < N) {
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
          if constexpr (kStage)
            *(uint2*)&stg[(lane & 15) * MP + m] = pack4_bf16(x.x, x.y, x.z, x.w);
          else
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
            const float4 rr = unpack4_bf16(rr16[r][t]);
            h = make_float4(h.x + rr.x, h.y + rr.y, h.z + rr.z, h.w + rr.w);
          }
          if (e.out) *(float4*)&e.out[(size_t)n * kH + m] = h;
          if constexpr (kStage)
            *(uint2*)&stg[(lane & 15) * MP + m] = pack4_bf16(h.x, h.y, h.z, h.w);
          else
            *(uint2*)&e.out16[(size_t)n * e.ld16 + m] = pack4_bf16(h.x, h.y, h.z, h.w);
        }
      }
      }
      if constexpr (kStage) {
        // the wave's 16 staged rows -> global, 16-byte chunks, rows contiguous
        __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
        __builtin_amdgcn_wave_barrier();
        constexpr int CH = M / 8;
        const uint32_t n0 = nb + r * 16;
        if (e.out16) {  // (null: rows only staged, for the fused projection)
#pragma unroll
          for (int c = lane; c < 16 * CH; c += 64) {
            const int row = c / CH, ch = c - row * CH;
            if (n0 + row < N)
              *(uint4*)&e.out16[(size_t)(n0 + row) * e.ld16 + ch * 8] = *(const uint4*)&stg[row * MP + ch * 8];
          }
        }
        if constexpr (FU) {
          // U of the staged rows: D = Wu H^T (2 tiles of 16 channels, 4 k-steps)
          f32x4 ua[kC / 16];
#pragma unroll
          for (int t2 = 0; t2 < kC / 16; t2++) ua[t2] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
          for (int kk = 0; kk < kH / 32; kk++) {
            const bf16x8 hb = *(const bf16x8*)&stg[(lane & 15) * MP + kk * 32 + kq];
#pragma unroll
            for (int t2 = 0; t2 < kC / 16; t2++) {
              const bf16x8 wa = *(const bf16x8*)&wul[(t2 * 16 + (lane & 15)) * KU + kk * 32 + kq];
              ua[t2] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa, hb, ua[t2], 0, 0, 0);
            }
          }
          const uint32_t nu = n0 + (lane & 15);
          if (nu < N) {
#pragma unroll
            for (int t2 = 0; t2 < kC / 16; t2++)
              *(uint2*)&e.u_out[(size_t)nu * kC + t2 * 16 + mq] =
                  pack4_bf16(ua[t2][0], ua[t2][1], ua[t2][2], ua[t2][3]);
          }
        }
        __builtin_amdgcn_s_waitcnt(0xc07f);
        __builtin_amdgcn_wave_barrier();
      }
    }
  }
#undef MP_LOAD_TILE
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

// g[r * gld + coff + k] += sum_chunks part[chunk][r][k] for r < rows, k < cols
// of a partial laid out [chunk][Mp][Kp] (the padded GEMM shape).
__global__ void k_reduce_rect(const float* __restrict__ part, uint32_t nchunk, int Mp, int Kp, int rows, int cols,
                              float* __restrict__ g, int gld, int coff) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * cols) return;
  const int r = i / cols, k = i % cols;
  float s = 0.f;
  for (uint32_t c = 0; c < nchunk; c++) s += part[((size_t)c * Mp + r) * Kp + k];
  g[(size_t)r * gld + coff + k] += s;
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
  return make_uint4(pk_bf16(f[0], f[1]), pk_bf16(f[2], f[3]), pk_bf16(f[4], f[5]), pk_bf16(f[6], f[7]));
}

// Edge weights rounded to bf16: the operand precision of the fused layer's
// aggregation MFMA, used by the training kernels as well so that training
// and inference evaluate the same function.
__device__ inline void edge_weights16(const float4 pi, const float4 pj, float& w, float& gx, float& gy) {
  edge_weights(pi, pj, w, gx, gy);
  w = bf2f(bf16_bits(w));
  gx = bf2f(bf16_bits(gx));
  gy = bf2f(bf16_bits(gy));
}

// Aggregation forward, one node per lane (consecutive nodes on consecutive
// lanes, so the slot-major edge lists load coalesced):
//   A[i][c] = sum_e w U[j][c], A[i][32+c] = sum_e gx U[j][c], A[i][64+c] = sum_e gy U[j][c]
__global__ __launch_bounds__(256) void k_agg_fwd(uint32_t n, uint32_t stride, const uint32_t* __restrict__ deg,
                                                 const uint32_t* __restrict__ col, const float4* __restrict__ npos,
                                                 const bf16_t* __restrict__ U, bf16_t* __restrict__ A) {
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
    edge_weights16(pi, pj, w, gx, gy);
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
  uint4* g = (uint4*)(A + (size_t)i * (3 * kC));
#pragma unroll
  for (int q = 0; q < kC / 8; q++) {
    g[q] = pack8_bf16(a0 + 8 * q);
    g[kC / 8 + q] = pack8_bf16(ax + 8 * q);
    g[2 * (kC / 8) + q] = pack8_bf16(ay + 8 * q);
  }
}

// One message-passing layer of the inference forward, fused into one kernel:
//   A  = [sum_e w U_j, sum_e gx U_j, sum_e gy U_j] over the edges of each node
//   H' = H + SiLU([H, A] Wo^T + bo)
//   then U' = H' Wu'^T (OUT == LO_U: the next layer's messages; H' stored) or
//   Y = H' Wd^T + bd (OUT == LO_DEC: the decoder; H' is not stored).
// A wave owns 16 nodes.  The aggregation is a batch of 16 small products on
// v_mfma_f32_4x4x4_16b_bf16 (block b = lanes 4b..4b+3 = one node): per batch
// of 4 edges, A = the 3 edge coefficients (w, gx, gy; row 3 zero) x 4 edges
// and B = 4 edges x 4 channels, one MFMA per channel offset m, where lane
// 4b + c loads the 16-byte slice (channels 8c..8c+7) of each neighbour's U
// row and computes the weights of edge c (exchanged through the wave's LDS
// stage).  The packed sums move to the dense GEMM's operand layout with
// ds_bpermute and feed its three message k-steps directly.  Waves are
// persistent and XCD-major: the waves of XCD x walk the x-th eighth of the
// 16-node tiles.
enum LayerOut : int { LO_U = 0, LO_DEC = 1 };
constexpr int kLayerThreads = 1024;

// FIRST: the layer's input H = SiLU(X0 We^T + be) is recomputed from the
// node features X0 (the encoder kernel stores only U), so the hidden state
// of the encoder never goes through memory.
template <int OUT, bool FIRST>
__global__ __launch_bounds__(kLayerThreads) void k_layer(
    uint32_t N, const bf16_t* __restrict__ H, const uint32_t* __restrict__ deg, const uint32_t* __restrict__ col,
    uint32_t stride, const float4* __restrict__ npos, const bf16_t* __restrict__ U, const bf16_t* __restrict__ Wo,
    const float* __restrict__ bo, const bf16_t* __restrict__ Wp, const float* __restrict__ bp,
    bf16_t* __restrict__ Hout, bf16_t* __restrict__ Uout, float* __restrict__ Y, const bf16_t* __restrict__ X0,
    const bf16_t* __restrict__ We, const float* __restrict__ be) {
  constexpr int K = kG, M = kH, KP = K + 8, T = M / 16;
  constexpr int kWaves = kLayerThreads / 64;
  constexpr int MP = M + 8;
  constexpr int PR = OUT == LO_U ? kC : kOutP;  // projection rows
  constexpr int KU = kH + 8;
  constexpr int KE = kK0 + 8;
  __shared__ __attribute__((aligned(16))) bf16_t wl[M * KP];
  __shared__ __attribute__((aligned(16))) bf16_t ostage[kWaves * 16 * MP];
  __shared__ __attribute__((aligned(16))) bf16_t wpl[PR * KU];
  __shared__ __attribute__((aligned(16))) bf16_t wel[FIRST ? kH * KE : 8];
  // biases in LDS: read per tile rather than held in registers across the tile loop
  __shared__ __attribute__((aligned(16))) float bl[kH];
  __shared__ __attribute__((aligned(16))) float bel[FIRST ? kH : 4];
  if (threadIdx.x < kH) bl[threadIdx.x] = bo[threadIdx.x];
  if constexpr (FIRST) {
    if (threadIdx.x < kH) bel[threadIdx.x] = be[threadIdx.x];
    for (int i = threadIdx.x * 8; i < kH * kK0; i += kLayerThreads * 8) {
      const int r = i / kK0, c = i - r * kK0;
      *(uint4*)&wel[r * KE + c] = *(const uint4*)&We[i];
    }
  }
  for (int i = threadIdx.x * 8; i < PR * kH; i += kLayerThreads * 8) {
    const int r = i / kH, c = i - r * kH;
    *(uint4*)&wpl[r * KU + c] = *(const uint4*)&Wp[i];
  }
  for (int i = threadIdx.x * 8; i < M * K; i += kLayerThreads * 8) {
    const int r = i / K, c = i - r * K;
    *(uint4*)&wl[r * KP + c] = *(const uint4*)&Wo[i];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int q = lane & 15, kg = lane >> 4;
  const int kq = kg * 8, mq = kg * 4;
  bf16_t* stg = ostage + wave * 16 * MP;
  bf16_t* ewt = stg;  // while aggregating: edge weights [16 nodes][4 rows][4 edges] bf16
  const uint32_t ntile = (N + 15) / 16;
  const uint32_t xcd = blockIdx.x % kXcds, bx = blockIdx.x / kXcds, nbx = gridDim.x / kXcds;
  const uint32_t t_lo = (uint32_t)((uint64_t)ntile * xcd / kXcds);
  const uint32_t t_hi = (uint32_t)((uint64_t)ntile * (xcd + 1) / kXcds);
  for (uint32_t tile = t_lo + bx * kWaves + wave; tile < t_hi; tile += nbx * kWaves) {
    const uint32_t n0 = tile * 16, n = n0 + q;
    const bool valid = n < N;
    const uint32_t nn = valid ? n : N - 1;  // clamped row for the loads of a partial tile
    // --- aggregation: lane = 4 * node + c (node an, channel block / coefficient row c)
    const int an = lane >> 2, ac = lane & 3;
    const uint32_t na = min(n0 + an, N - 1);
    const bool va = n0 + an < N;
    const float4 pia = *(const float4*)((const char*)npos + (na << 4));
    const uint32_t da = va ? deg[na] : 0u;
    uint32_t dm = da;
#pragma unroll
    for (int o = 4; o < 64; o <<= 1) dm = max(dm, (uint32_t)__shfl_xor((int)dm, o, 64));
    dm = __builtin_amdgcn_readfirstlane(dm);
    f32x4 ag[8];  // ag[m][r]: sum of coefficient r (w, gx, gy, -) times channel 8 c + m
#pragma unroll
    for (int m = 0; m < 8; m++) ag[m] = f32x4{0.f, 0.f, 0.f, 0.f};
    // Software pipeline over batches of 4 edges: the neighbour ids of batch
    // b + 2 and the message slices and positions of batch b + 1 are in flight
    // while batch b is reduced.  Slots past the node's degree load slot 63
    // (in bounds) and use the node itself.
    uint32_t jr[4];  // raw slot contents of the next batch (selected when used)
    auto load_ids = [&](uint32_t kb) {
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const uint32_t slot = min(kb + e, kMaxDeg - 1);
        jr[e] = *(const uint32_t*)((const char*)col + ((slot * stride + na) << 2));
      }
    };
    uint4 uv[4];  // channels 8c..8c+7 of the 4 neighbours' messages
    float4 pc;    // position of the lane's edge (edge c of the batch)
    auto load_batch = [&](uint32_t kb) {
      uint32_t j[4];
#pragma unroll
      for (int e = 0; e < 4; e++) j[e] = kb + e < da ? jr[e] : na;
#pragma unroll
      for (int e = 0; e < 4; e++) uv[e] = *(const uint4*)((const char*)U + (j[e] * (kC * 2u) + ac * 16u));
      const uint32_t jc = ac == 0 ? j[0] : ac == 1 ? j[1] : ac == 2 ? j[2] : j[3];
      pc = *(const float4*)((const char*)npos + (jc << 4));
    };
    if (dm > 0) {
      load_ids(0);
      load_batch(0);
      load_ids(4);
    }
    for (uint32_t k0 = 0; k0 < dm; k0 += 4) {
      const uint4 cu[4] = {uv[0], uv[1], uv[2], uv[3]};
      const float4 cp = pc;
      // (unconditional: past the last batch these are in-bounds loads of the node's own row)
      load_batch(k0 + 4);
      load_ids(k0 + 8);
      // weights of edge c, as bf16, into the wave's [node][row][edge] table
      float w, gx, gy;
      edge_weights(pia, cp, w, gx, gy);
      if (k0 + ac >= da) w = gx = gy = 0.f;
      ewt[(an * 4 + 0) * 4 + ac] = bf16_bits(w);
      ewt[(an * 4 + 1) * 4 + ac] = bf16_bits(gx);
      ewt[(an * 4 + 2) * 4 + ac] = bf16_bits(gy);
      __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
      __builtin_amdgcn_wave_barrier();
      // A operand: row c = coefficient c over the 4 edges (row 3 zero)
      const s16x4 ea = ac < 3 ? *(const s16x4*)&ewt[(an * 4 + ac) * 4] : s16x4{0, 0, 0, 0};
      // B operands: channel 8c + m of the 4 edges; D += A B per node (16 blocks)
      const uint32_t u0[4] = {cu[0].x, cu[0].y, cu[0].z, cu[0].w}, u1[4] = {cu[1].x, cu[1].y, cu[1].z, cu[1].w};
      const uint32_t u2[4] = {cu[2].x, cu[2].y, cu[2].z, cu[2].w}, u3[4] = {cu[3].x, cu[3].y, cu[3].z, cu[3].w};
#pragma unroll
      for (int dw = 0; dw < 4; dw++) {
        const uint2 lo = make_uint2(__builtin_amdgcn_perm(u1[dw], u0[dw], 0x05040100u),
                                    __builtin_amdgcn_perm(u3[dw], u2[dw], 0x05040100u));
        const uint2 hi = make_uint2(__builtin_amdgcn_perm(u1[dw], u0[dw], 0x07060302u),
                                    __builtin_amdgcn_perm(u3[dw], u2[dw], 0x07060302u));
        ag[2 * dw] = __builtin_amdgcn_mfma_f32_4x4x4bf16_1k(ea, __builtin_bit_cast(s16x4, lo), ag[2 * dw], 0, 0, 0);
        ag[2 * dw + 1] =
            __builtin_amdgcn_mfma_f32_4x4x4bf16_1k(ea, __builtin_bit_cast(s16x4, hi), ag[2 * dw + 1], 0, 0, 0);
      }
      __builtin_amdgcn_wave_barrier();
    }
    bf16x8 xh[kH / 32];  // the GEMM's H operand
    if constexpr (FIRST) {
      // H = SiLU(X0 We^T + be) for this tile, staged as bf16 rows (also the residual)
      const bf16x8 xb0 = *(const bf16x8*)(X0 + (size_t)nn * kK0 + kq);
#pragma unroll
      for (int t = 0; t < T; t++) {
        const bf16x8 wa = *(const bf16x8*)&wel[(t * 16 + q) * KE + kq];
        const f32x4 z0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa, xb0, f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
        const int m = t * 16 + mq;
        const float4 bb = *(const float4*)&bel[m];
        *(uint2*)&stg[q * MP + m] =
            pack4_bf16(silu(z0[0] + bb.x), silu(z0[1] + bb.y), silu(z0[2] + bb.z), silu(z0[3] + bb.w));
      }
      __builtin_amdgcn_s_waitcnt(0xc07f);
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int kk = 0; kk < kH / 32; kk++) xh[kk] = *(const bf16x8*)&stg[q * MP + kk * 32 + kq];
    } else {
#pragma unroll
      for (int kk = 0; kk < kH / 32; kk++) xh[kk] = *(const bf16x8*)(H + (size_t)nn * kH + kk * 32 + kq);
    }
    // packed sums, then lane 4 q + g -> lane 16 g + q (the GEMM operand layout)
    uint32_t pk[3][4];
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
      for (int dw = 0; dw < 4; dw++) pk[r][dw] = pk_bf16(ag[2 * dw][r], ag[2 * dw + 1][r]);
    const int src = (4 * q + kg) << 2;
    bf16x8 xa[3];
#pragma unroll
    for (int r = 0; r < 3; r++) {
      const uint4 v = make_uint4(__builtin_amdgcn_ds_bpermute(src, pk[r][0]), __builtin_amdgcn_ds_bpermute(src, pk[r][1]),
                                 __builtin_amdgcn_ds_bpermute(src, pk[r][2]), __builtin_amdgcn_ds_bpermute(src, pk[r][3]));
      xa[r] = __builtin_bit_cast(bf16x8, v);
    }
    f32x4 acc[T];
#pragma unroll
    for (int t = 0; t < T; t++) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kk = 0; kk < K / 32; kk++) {
      const bf16x8 xb = kk < kH / 32 ? xh[kk] : xa[kk - kH / 32];
#pragma unroll
      for (int t = 0; t < T; t++) {
        const bf16x8 wa = *(const bf16x8*)&wl[(t * 16 + q) * KP + kk * 32 + kq];
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa, xb, acc[t], 0, 0, 0);
      }
      __builtin_amdgcn_sched_barrier(0);
    }
    // H' = H + SiLU(z) into the stage (rows of this tile)
#pragma unroll
    for (int t = 0; t < T; t++) {
      const int m = t * 16 + mq;
      const float4 rr = unpack4_bf16(FIRST ? *(const uint2*)&stg[q * MP + m] : *(const uint2*)&H[(size_t)nn * kH + m]);
      const float4 bb = *(const float4*)&bl[m];
      const float4 z = make_float4(acc[t][0] + bb.x, acc[t][1] + bb.y, acc[t][2] + bb.z, acc[t][3] + bb.w);
      *(uint2*)&stg[q * MP + m] =
          pack4_bf16(rr.x + silu(z.x), rr.y + silu(z.y), rr.z + silu(z.z), rr.w + silu(z.w));
    }
    __builtin_amdgcn_s_waitcnt(0xc07f);
    __builtin_amdgcn_wave_barrier();
    if constexpr (OUT == LO_U) {
      constexpr int CH = M / 8;
#pragma unroll
      for (int c = lane; c < 16 * CH; c += 64) {
        const int row = c / CH, ch = c - row * CH;
        if (n0 + row < N) *(uint4*)&Hout[(size_t)(n0 + row) * kH + ch * 8] = *(const uint4*)&stg[row * MP + ch * 8];
      }
    }
    // projection of the staged rows: D = Wp H'^T
    f32x4 pa[PR / 16];
#pragma unroll
    for (int t2 = 0; t2 < PR / 16; t2++) pa[t2] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kk = 0; kk < kH / 32; kk++) {
      const bf16x8 hb = *(const bf16x8*)&stg[q * MP + kk * 32 + kq];
#pragma unroll
      for (int t2 = 0; t2 < PR / 16; t2++) {
        const bf16x8 wa = *(const bf16x8*)&wpl[(t2 * 16 + q) * KU + kk * 32 + kq];
        pa[t2] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa, hb, pa[t2], 0, 0, 0);
      }
    }
    if (valid) {
      if constexpr (OUT == LO_U) {
#pragma unroll
        for (int t2 = 0; t2 < PR / 16; t2++)
          *(uint2*)&Uout[(size_t)n * kC + t2 * 16 + mq] = pack4_bf16(pa[t2][0], pa[t2][1], pa[t2][2], pa[t2][3]);
      } else {
        float4 y = make_float4(pa[0][0], pa[0][1], pa[0][2], pa[0][3]);
        y.x += mq < kOut ? bp[mq] : 0.f;
        y.y += mq + 1 < kOut ? bp[mq + 1] : 0.f;
        y.z += mq + 2 < kOut ? bp[mq + 2] : 0.f;
        y.w += mq + 3 < kOut ? bp[mq + 3] : 0.f;
        *(float4*)&Y[(size_t)n * kOutP + mq] = y;
      }
    }
    __builtin_amdgcn_s_waitcnt(0xc07f);
    __builtin_amdgcn_wave_barrier();
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
    edge_weights16(pj, pi, w, gx, gy);  // (gx, gy)_ji = -(gx, gy)_ij
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
template <int K, int M, int KS = K, int ROWS = (M > 32 ? 1 : 4)>
void gemm2(const bf16_t* X, int ldx, const bf16_t* X2, int ldx2, const bf16_t* W, uint32_t N, const GemmEpi& e,
           hipStream_t st) {
  constexpr int TH = gemm_threads<K, M>();
  const uint32_t tiles = blocks(N, (TH / 64) * ROWS * 16);
  const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(tiles, gemm_blocks_max<K, M>()));
  switch (e.mode) {
    case EPI_F32:
      hipLaunchKernelGGL((k_gemm<K, M, ROWS, EPI_F32, KS, TH>), dim3(grid), dim3(TH), 0, st, X, ldx, X2, ldx2, W, N,
                         e);
      break;
    case EPI_F32_ACC:
      hipLaunchKernelGGL((k_gemm<K, M, ROWS, EPI_F32_ACC, KS, TH>), dim3(grid), dim3(TH), 0, st, X, ldx, X2, ldx2, W,
                         N, e);
      break;
    case EPI_BF16:
      hipLaunchKernelGGL((k_gemm<K, M, ROWS, EPI_BF16, KS, TH>), dim3(grid), dim3(TH), 0, st, X, ldx, X2, ldx2, W, N,
                         e);
      break;
    case EPI_SPLIT:
      hipLaunchKernelGGL((k_gemm<K, M, ROWS, EPI_SPLIT, KS, TH>), dim3(grid), dim3(TH), 0, st, X, ldx, X2, ldx2, W,
                         N, e);
      break;
    default:
      if constexpr (M == kH) {
        if (e.wu) {
          hipLaunchKernelGGL((k_gemm<K, M, ROWS, EPI_SILU_RES, KS, TH, true>), dim3(grid), dim3(TH), 0, st, X, ldx,
                             X2, ldx2, W, N, e);
          break;
        }
      }
      hipLaunchKernelGGL((k_gemm<K, M, ROWS, EPI_SILU_RES, KS, TH>), dim3(grid), dim3(TH), 0, st, X, ldx, X2, ldx2,
                         W, N, e);
      break;
  }
}

template <int K, int M>
void gemm(const bf16_t* X, int ldx, const bf16_t* W, uint32_t N, const GemmEpi& e, hipStream_t st) {
  gemm2<K, M, K>(X, ldx, X, ldx, W, N, e, st);
}

GemmEpi epi(int mode) {
  GemmEpi e;
  memset(&e, 0, sizeof(e));
  e.mode = mode;
  return e;
}

// g[r * gld + coff + k] += (dY^T X)[r][k] over all N nodes, r < rows (of the
// M-row padded product), k < K
template <int K>
void gemm_dw(const bf16_t* dY, int ldy, const bf16_t* X, int ldx, uint32_t N, int M, int rows, float* part, float* g,
             int gld, int coff, hipStream_t st) {
  const uint32_t nch = blocks(N, kDwChunk);
  hipLaunchKernelGGL((k_gemm_dw<K>), dim3(M / 16, nch), dim3(256), 0, st, dY, ldy, X, ldx, N, M, part);
  hipLaunchKernelGGL(k_reduce_rect, dim3(blocks((size_t)rows * K, 256)), dim3(256), 0, st, part, nch, M, K, rows, K,
                     g, gld, coff);
}

void bias_grad(const bf16_t* dY, int ldy, uint32_t N, int M, float* part, float* g, hipStream_t st) {
  const uint32_t nch = blocks(N, kDwChunk);
  hipLaunchKernelGGL(k_colsum, dim3(nch), dim3(256), 0, st, dY, ldy, N, M, part);
  hipLaunchKernelGGL(k_reduce_rect, dim3(blocks(M, 256)), dim3(256), 0, st, part, nch, 1, M, 1, M, g, M, 0);
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
  for (int l = 0; l <= kLayers && ok; l++) ok = al((void**)&a.H16[l], N * kH * 2);
  for (int l = 0; l < kLayers && ok; l++) ok = al((void**)&a.A16[l], N * 3 * kC * 2) && al((void**)&a.U[l], N * kC * 2);
  if (train) {
    ok = ok && al((void**)&a.dH, N * kH * 4) && al((void**)&a.dZ16, N * kH * 2) && al((void**)&a.dU16, N * kC * 2) &&
         al((void**)&a.dA16, N * 3 * kC * 2);
    for (int l = 0; l <= kLayers && ok; l++) ok = al((void**)&a.Z[l], N * kH * 4) && al((void**)&a.Hf[l], N * kH * 4);
    a.part_cap = (size_t)blocks(N, kDwChunk) * kH * kH;
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
  a.dZ16 = a.dU16 = a.dA16 = nullptr;
  a.part = nullptr;
}

void mp_forward_layers(const MpNet& net, MpActs& a, const MpGraph& g, uint32_t n, hipStream_t st) {
  if (n == 0) return;
  static_assert(kLayers == 3, "fused inference forward: first, middle and last layer");
  const uint32_t grid = std::min<uint32_t>(256, xcd_grid(blocks(blocks(n, 16), kLayerThreads / 64)));
  const bf16_t* we = net.t[net.We].w16;
  const float* be = net.w + net.t[net.be].off;
  auto bo = [&](int l) { return net.w + net.t[net.bo[l]].off; };
  hipLaunchKernelGGL((k_layer<LO_U, true>), dim3(grid), dim3(kLayerThreads), 0, st, n, nullptr, g.deg, g.col, g.cap,
                     g.npos, a.U[0], net.t[net.Wo[0]].w16, bo(0), net.t[net.Wu[1]].w16, nullptr, a.H16[1], a.U[1],
                     nullptr, a.X0, we, be);
  hipLaunchKernelGGL((k_layer<LO_U, false>), dim3(grid), dim3(kLayerThreads), 0, st, n, a.H16[1], g.deg, g.col,
                     g.cap, g.npos, a.U[1], net.t[net.Wo[1]].w16, bo(1), net.t[net.Wu[2]].w16, nullptr, a.H16[2],
                     a.U[2], nullptr, nullptr, nullptr, nullptr);
  hipLaunchKernelGGL((k_layer<LO_DEC, false>), dim3(grid), dim3(kLayerThreads), 0, st, n, a.H16[2], g.deg, g.col,
                     g.cap, g.npos, a.U[2], net.t[net.Wo[2]].w16, bo(2), net.t[net.Wd].w16,
                     net.w + net.t[net.bd].off, nullptr, nullptr, a.Y, nullptr, nullptr, nullptr);
}

void mp_forward(const MpNet& net, MpActs& a, const MpGraph& g, uint32_t n, bool train, hipStream_t st) {
  if (n == 0) return;
  // encoder: H0 = SiLU(X0 We^T + be)
  GemmEpi e = epi(EPI_SILU_RES);
  e.bias = net.w + net.t[net.be].off;
  e.nbias = kH;
  e.zout = train ? a.Z[0] : nullptr;
  e.out = train ? a.Hf[0] : nullptr;
  e.out16 = a.H16[0];
  e.ld16 = kH;
  e.wu = net.t[net.Wu[0]].w16;  // U_0 = H_0 Wu_0^T from the staged rows
  e.u_out = a.U[0];
  if (!train) {
    // the encoder stores only U_0; the first fused layer recomputes H_0 from
    // X0, the last one ends in the decoder
    e.out16 = nullptr;
    gemm<kK0, kH>(a.X0, kK0, net.t[net.We].w16, n, e, st);
    mp_forward_layers(net, a, g, n, st);
    return;
  }
  gemm<kK0, kH>(a.X0, kK0, net.t[net.We].w16, n, e, st);
  for (int l = 0; l < kLayers; l++) {
    // A = aggregate(U)
    hipLaunchKernelGGL(k_agg_fwd, dim3(xcd_grid(blocks(n, 256))), dim3(256), 0, st, n, g.cap, g.deg, g.col, g.npos,
                       a.U[l], a.A16[l]);
    // H_{l+1} = H_l + SiLU([H, A] Wo^T + bo)
    GemmEpi eo = epi(EPI_SILU_RES);
    eo.bias = net.w + net.t[net.bo[l]].off;
    eo.nbias = kH;
    eo.zout = train ? a.Z[l + 1] : nullptr;
    eo.res = train ? a.Hf[l] : nullptr;
    eo.res16 = train ? nullptr : a.H16[l];
    eo.out = train ? a.Hf[l + 1] : nullptr;
    eo.out16 = a.H16[l + 1];
    eo.ld16 = kH;
    if (l + 1 < kLayers) {  // U_{l+1} = H_{l+1} Wu_{l+1}^T
      eo.wu = net.t[net.Wu[l + 1]].w16;
      eo.u_out = a.U[l + 1];
    }
    gemm2<kG, kH, kH>(a.H16[l], kH, a.A16[l], 3 * kC, net.t[net.Wo[l]].w16, n, eo, st);
  }
  GemmEpi ed = epi(EPI_F32);
  ed.bias = net.w + net.t[net.bd].off;
  ed.nbias = kOut;
  ed.out = a.Y;
  ed.ldo = kOutP;
  gemm<kH, kOutP>(a.H16[kLayers], kH, net.t[net.Wd].w16, n, ed, st);
}

mfree_status mp_graph_reserve(MpGraph& g, uint32_t n, std::string& msg) {
  if (n <= g.cap) return MFREE_OK;
  if ((uint64_t)n * kMaxDeg * 4 >= (1ull << 32)) {  // the layer kernels address edge slots with 32-bit offsets
    msg = "learned-step graph limited to 16M nodes";
    return MFREE_ERR_CAPACITY;
  }
  mp_graph_free(g);
  auto al = [&](void** p, size_t bytes) -> bool {
    if (hipMalloc(p, std::max<size_t>(bytes, 16)) != hipSuccess) return false;
    g.allocs.push_back(*p);
    return true;
  };
  const size_t N = n;
  if (!al((void**)&g.deg, N * 4) || !al((void**)&g.col, N * kMaxDeg * 4) || !al((void**)&g.npos, N * 16) ||
      !al((void**)&g.dpos, N * 32) || !al((void**)&g.wsum, N * 4)) {
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
  hipLaunchKernelGGL(k_mse, dim3(blocks(n, 256)), dim3(256), 0, st, n, a.Y, T, a.dU16, loss_out);
  // decoder: dWd = dY^T H3 (rows 0..kOut-1 of the padded 16), dbd = sum dY, dH3 = dY Wd
  gemm_dw<kH>(a.dU16, kC, a.H16[kLayers], kH, n, 16, kOut, a.part, net.g + net.t[net.Wd].off, kH, 0, st);
  bias_grad(a.dU16, kC, n, kOut, a.part, net.g + net.t[net.bd].off, st);
  GemmEpi e = epi(EPI_F32);
  e.out = a.dH;
  e.ldo = kH;
  gemm<kC, kH>(a.dU16, kC, net.t[net.Wd].wt16, n, e, st);
  for (int l = kLayers - 1; l >= 0; l--) {
    // dZ = dH * silu'(Z_{l+1}); the residual keeps dH for H_l
    hipLaunchKernelGGL(k_silu_bwd, dim3(blocks((size_t)n * kH, 256)), dim3(256), 0, st, n * kH, a.dH, a.Z[l + 1],
                       a.dZ16, kH);
    // dWo = dZ^T [H_l, A_l], dbo = sum dZ
    float* gWo = net.g + net.t[net.Wo[l]].off;
    gemm_dw<kH>(a.dZ16, kH, a.H16[l], kH, n, kH, kH, a.part, gWo, kG, 0, st);
    gemm_dw<3 * kC>(a.dZ16, kH, a.A16[l], 3 * kC, n, kH, kH, a.part, gWo, kG, kH, st);
    bias_grad(a.dZ16, kH, n, kH, a.part, net.g + net.t[net.bo[l]].off, st);
    // [dH_l += dG[:, :128], dA = dG[:, 128:224]] with dG = dZ Wo
    GemmEpi eg = epi(EPI_SPLIT);
    eg.out = a.dH;
    eg.out16 = a.dA16;
    gemm<kH, kG>(a.dZ16, kH, net.t[net.Wo[l]].wt16, n, eg, st);
    // dU = aggregate^T(dA)
    hipLaunchKernelGGL(k_agg_bwd, dim3(xcd_grid(blocks(n, 256))), dim3(256), 0, st, n, g.cap, g.deg, g.col, g.npos,
                       a.dA16, a.dU16, kC);
    // dWu = dU^T H_l, dH_l += dU Wu
    gemm_dw<kH>(a.dU16, kC, a.H16[l], kH, n, kC, kC, a.part, net.g + net.t[net.Wu[l]].off, kH, 0, st);
    GemmEpi eu = epi(EPI_F32_ACC);
    eu.out = a.dH;
    eu.ldo = kH;
    gemm<kC, kH>(a.dU16, kC, net.t[net.Wu[l]].wt16, n, eu, st);
  }
  // encoder: dZ0 = dH0 * silu'(Z0); dWe = dZ0^T X0 (first kNF columns), dbe
  hipLaunchKernelGGL(k_silu_bwd, dim3(blocks((size_t)n * kH, 256)), dim3(256), 0, st, n * kH, a.dH, a.Z[0], a.dZ16,
                     kH);
  const uint32_t nch = blocks(n, kDwChunk);
  hipLaunchKernelGGL((k_gemm_dw<kK0>), dim3(kH / 16, nch), dim3(256), 0, st, a.dZ16, kH, a.X0, kK0, n, kH, a.part);
  hipLaunchKernelGGL(k_reduce_rect, dim3(blocks((size_t)kH * kNF, 256)), dim3(256), 0, st, a.part, nch, kH, kK0, kH,
                     kNF, net.g + net.t[net.We].off, kNF, 0);
  bias_grad(a.dZ16, kH, n, kH, a.part, net.g + net.t[net.be].off, st);
}

void mp_adam(MpNet& net, float lr, float beta1, float beta2, float eps, hipStream_t st) {
  net.adam_t++;
  const float c1 = 1.f - powf(beta1, (float)net.adam_t), c2 = 1.f - powf(beta2, (float)net.adam_t);
  hipLaunchKernelGGL(k_adam, dim3(blocks(net.nparam, 256)), dim3(256), 0, st, net.nparam, net.w, net.g, net.m1,
                     net.m2, lr, beta1, beta2, eps, c1, c2);
  mp_net_sync16(net, st);
}

}  // namespace mfree
