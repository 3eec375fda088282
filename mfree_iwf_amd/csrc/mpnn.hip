// Message-passing surrogate: forward and backward kernels on gfx950.
//
// Dense layers: k_gemm computes out[n][m] = sum_k X[n][k] W[m][k] with
// v_mfma_f32_16x16x32_bf16 (a wave owns 16 nodes and every 16-wide output
// tile; operands are 8 contiguous bf16 per lane: A row = lane%16, B column =
// lane%16, result column = lane%16 / rows 4*(lane/16)+v) and fused
// epilogues (bias, SiLU, residual, FP32 / BF16 stores).  Weight gradients
// (k_gemm_dw) use the same instruction with the node axis as the reduction
// axis and per-chunk partial sums.  The aggregation over the neighbour
// graph is a gather per node; its backward is the same gather with the
// signs of the antisymmetric gradient weights flipped (the graph and w are
// symmetric, gx and gy antisymmetric).
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
};

struct GemmEpi {
  int mode;
  const float* bias;            // nbias entries (columns beyond get no bias)
  int nbias;
  float* out;
  int ldo;
  bf16_t* out16;
  int ld16;
  float* zout;
  const float* res;
};

template <int K, int M>
__global__ __launch_bounds__(256) void k_gemm(const bf16_t* __restrict__ X, int ldx, const bf16_t* __restrict__ W,
                                              uint32_t N, GemmEpi e) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t nb = blockIdx.x * 64u + wave * 16u;
  if (nb >= N) return;
  const uint32_t row = nb + (lane & 15);
  const int kq = (lane >> 4) * 8;
  constexpr int T = M / 16;
  f32x4 acc[T];
#pragma unroll
  for (int t = 0; t < T; t++) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int k0 = 0; k0 < K; k0 += 32) {
    bf16x8 a;
    if (row < N)
      a = *(const bf16x8*)(X + (size_t)row * ldx + k0 + kq);
    else
      a = bf16x8{};
#pragma unroll
    for (int t = 0; t < T; t++) {
      const bf16x8 b = *(const bf16x8*)(W + (size_t)(t * 16 + (lane & 15)) * K + k0 + kq);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
    }
  }
#pragma unroll
  for (int t = 0; t < T; t++) {
    const int m = t * 16 + (lane & 15);
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const uint32_t n = nb + (lane >> 4) * 4 + v;
      if (n >= N) continue;
      float r = acc[t][v];
      const float bm = (e.bias && m < e.nbias) ? e.bias[m] : 0.f;
      if (e.mode == EPI_F32) {
        e.out[(size_t)n * e.ldo + m] = r + bm;
      } else if (e.mode == EPI_F32_ACC) {
        e.out[(size_t)n * e.ldo + m] += r;
      } else if (e.mode == EPI_BF16) {
        e.out16[(size_t)n * e.ld16 + m] = f2bf(r + bm);
      } else {
        const float z = r + bm;
        if (e.zout) e.zout[(size_t)n * kH + m] = z;
        const float h = (e.res ? e.res[(size_t)n * kH + m] : 0.f) + silu(z);
        e.out[(size_t)n * kH + m] = h;
        e.out16[(size_t)n * e.ld16 + m] = f2bf(h);
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

// Aggregation forward: for node i and channel c,
//   G[i][128+c] = sum_e w U[j][c], G[i][160+c] = sum_e gx U[j][c], G[i][192+c] = sum_e gy U[j][c]
// 32 lanes per node.
__global__ __launch_bounds__(256) void k_agg_fwd(uint32_t n, const uint32_t* __restrict__ rowp,
                                                 const uint32_t* __restrict__ col, const float4* __restrict__ ew,
                                                 const bf16_t* __restrict__ U, bf16_t* __restrict__ G) {
  const uint32_t i = blockIdx.x * 8u + (threadIdx.x >> 5);
  const int c = threadIdx.x & 31;
  if (i >= n) return;
  float a0 = 0.f, ax = 0.f, ay = 0.f;
  const uint32_t e0 = rowp[i], e1 = rowp[i + 1];
  for (uint32_t e = e0; e < e1; e++) {
    const float4 w = ew[e];
    const float u = bf2f(U[(size_t)col[e] * kC + c]);
    a0 += w.x * u;
    ax += w.y * u;
    ay += w.z * u;
  }
  bf16_t* g = G + (size_t)i * kG + kH;
  g[c] = f2bf(a0);
  g[kC + c] = f2bf(ax);
  g[2 * kC + c] = f2bf(ay);
}

// Aggregation backward: dU[j][c] = sum_{i in N(j)} (w dA0_i - gx_ji dAx_i - gy_ji dAy_i)
// (w symmetric, gx/gy antisymmetric); dA = dG[:, 128:224].
__global__ __launch_bounds__(256) void k_agg_bwd(uint32_t n, const uint32_t* __restrict__ rowp,
                                                 const uint32_t* __restrict__ col, const float4* __restrict__ ew,
                                                 const float* __restrict__ dG, float* __restrict__ dU) {
  const uint32_t j = blockIdx.x * 8u + (threadIdx.x >> 5);
  const int c = threadIdx.x & 31;
  if (j >= n) return;
  float s = 0.f;
  const uint32_t e0 = rowp[j], e1 = rowp[j + 1];
  for (uint32_t e = e0; e < e1; e++) {
    const float4 w = ew[e];
    const float* d = dG + (size_t)col[e] * kG + kH;
    s += w.x * d[c] - w.y * d[kC + c] - w.z * d[2 * kC + c];
  }
  dU[(size_t)j * kC + c] = s;
}

// dZ = dH * silu'(Z) -> bf16 operand (ld 224 buffer, columns [0,128))
__global__ void k_silu_bwd(uint32_t nelem, const float* __restrict__ dH, const float* __restrict__ Z,
                           bf16_t* __restrict__ dZ16, int ld16) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nelem) return;
  const uint32_t n = idx / kH, m = idx % kH;
  dZ16[(size_t)n * ld16 + m] = f2bf(dH[idx] * silu_d(Z[idx]));
}

// f32 [N x cols] -> bf16 [N x ld16] (columns [0, cols))
__global__ void k_to_bf16(uint32_t N, int cols, const float* __restrict__ src, int lds, bf16_t* __restrict__ dst,
                          int ldd) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= N * (uint32_t)cols) return;
  const uint32_t n = idx / cols, m = idx % cols;
  dst[(size_t)n * ldd + m] = f2bf(src[(size_t)n * lds + m]);
}

// dH += dG[:, :128]  (residual + the H part of [H, A])
__global__ void k_add_dH(uint32_t nelem, const float* __restrict__ dG, float* __restrict__ dH) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nelem) return;
  const uint32_t n = idx / kH, m = idx % kH;
  dH[idx] += dG[(size_t)n * kG + m];
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

template <int K, int M>
void gemm(const bf16_t* X, int ldx, const bf16_t* W, uint32_t N, const GemmEpi& e, hipStream_t st) {
  hipLaunchKernelGGL((k_gemm<K, M>), dim3(blocks(N, 64)), dim3(256), 0, st, X, ldx, W, N, e);
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

mfree_status mp_acts_reserve(MpActs& a, uint32_t n, std::string& msg) {
  if (n <= a.cap) return MFREE_OK;
  mp_acts_free(a);
  auto al = [&](void** p, size_t bytes) -> bool {
    if (hipMalloc(p, std::max<size_t>(bytes, 16)) != hipSuccess) return false;
    a.allocs.push_back(*p);
    return true;
  };
  const size_t N = n;
  bool ok = al((void**)&a.X0, N * kK0 * 2) && al((void**)&a.Y, N * kOutP * 4) && al((void**)&a.dH, N * kH * 4) &&
            al((void**)&a.dZ16, N * kG * 2) && al((void**)&a.dG, N * kG * 4) && al((void**)&a.dA16, N * 96 * 2) &&
            al((void**)&a.dU, N * kC * 4);
  for (int l = 0; l <= kLayers && ok; l++)
    ok = al((void**)&a.Z[l], N * kH * 4) && al((void**)&a.Hf[l], N * kH * 4) && al((void**)&a.G[l], N * kG * 2);
  for (int l = 0; l < kLayers && ok; l++) ok = al((void**)&a.U[l], N * kC * 2);
  a.part_cap = (size_t)blocks(N, kDwChunk) * kH * kG;
  ok = ok && al((void**)&a.part, a.part_cap * 4);
  if (!ok) {
    mp_acts_free(a);
    msg = "hipMalloc failed (activations)";
    return MFREE_ERR_HIP;
  }
  a.cap = n;
  return MFREE_OK;
}

void mp_acts_free(MpActs& a) {
  for (void* p : a.allocs) (void)hipFree(p);
  a.allocs.clear();
  a.cap = 0;
}

void mp_forward(const MpNet& net, MpActs& a, const MpGraph& g, uint32_t n, bool train, hipStream_t st) {
  // encoder: H0 = SiLU(X0 We^T + be)
  GemmEpi e = epi(EPI_SILU_RES);
  e.bias = net.w + net.t[net.be].off;
  e.nbias = kH;
  e.zout = train ? a.Z[0] : nullptr;
  e.out = a.Hf[0];
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
    hipLaunchKernelGGL(k_agg_fwd, dim3(blocks(n, 8)), dim3(256), 0, st, n, g.rowp, g.col, g.ew, a.U[l], a.G[l]);
    // H_{l+1} = H_l + SiLU([H, A] Wo^T + bo)
    GemmEpi eo = epi(EPI_SILU_RES);
    eo.bias = net.w + net.t[net.bo[l]].off;
    eo.nbias = kH;
    eo.zout = train ? a.Z[l + 1] : nullptr;
    eo.res = a.Hf[l];
    eo.out = a.Hf[l + 1];
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
  const uint64_t E = (uint64_t)n * kEdgesPerNode;
  if (!al((void**)&g.deg, (N + 1) * 4) || !al((void**)&g.rowp, (N + 1) * 4) ||
      !al((void**)&g.tsum, ((N + 1 + kScanTile - 1) / kScanTile + 1) * 4) || !al((void**)&g.col, E * 4) ||
      !al((void**)&g.ew, E * 16) || !al((void**)&g.wsum, N * 4)) {
    mp_graph_free(g);
    msg = "hipMalloc failed (graph)";
    return MFREE_ERR_HIP;
  }
  g.ecap = E;
  g.cap = n;
  return MFREE_OK;
}

void mp_graph_free(MpGraph& g) {
  for (void* p : g.allocs) (void)hipFree(p);
  g.allocs.clear();
  g.cap = 0;
  g.ecap = 0;
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
    // dZ = dH * silu'(Z_{l+1}); residual keeps dH for H_l
    hipLaunchKernelGGL(k_silu_bwd, dim3(blocks((size_t)n * kH, 256)), dim3(256), 0, st, n * kH, a.dH, a.Z[l + 1],
                       a.dZ16, kG);
    // dWo = dZ^T G_l, dbo = sum dZ, dG = dZ Wo
    gemm_dw<kG>(a.dZ16, kG, a.G[l], kG, n, kH, a.part, net.g + net.t[net.Wo[l]].off, st);
    bias_grad(a.dZ16, kG, n, kH, a.part, net.g + net.t[net.bo[l]].off, st);
    GemmEpi eg = epi(EPI_F32);
    eg.out = a.dG;
    eg.ldo = kG;
    gemm<kH, kG>(a.dZ16, kG, net.t[net.Wo[l]].wt16, n, eg, st);
    // dH_l = dH_{l+1} + dG[:, :128]
    hipLaunchKernelGGL(k_add_dH, dim3(blocks((size_t)n * kH, 256)), dim3(256), 0, st, n * kH, a.dG, a.dH);
    // dU = aggregate^T(dA)
    hipLaunchKernelGGL(k_agg_bwd, dim3(blocks(n, 8)), dim3(256), 0, st, n, g.rowp, g.col, g.ew, a.dG, a.dU);
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks((size_t)n * kC, 256)), dim3(256), 0, st, n, kC, a.dU, kC, a.dZ16, kG);
    // dWu = dU^T H_l, dH_l += dU Wu
    gemm_dw<kH>(a.dZ16, kG, a.G[l], kG, n, kC, a.part, net.g + net.t[net.Wu[l]].off, st);
    GemmEpi eu = epi(EPI_F32_ACC);
    eu.out = a.dH;
    eu.ldo = kH;
    gemm<kC, kH>(a.dZ16, kG, net.t[net.Wu[l]].wt16, n, eu, st);
  }
  // encoder: dZ0 = dH0 * silu'(Z0); dWe = dZ0^T X0 (first 9 columns), dbe
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
