// Message-passing surrogate stepper (internal): a graph network over the SPH
// neighbour graph that predicts the correction of a 4-step state update.
//
// Architecture (hidden width H = 128, message width C = 32, L = 3 layers):
//   X0  = node features (NF = 16, zero-padded to 32)            [N x 32]  bf16
//   H0  = SiLU(X0 We^T + be)                                    [N x 128]
//   layer l:  U  = H Wu^T  (fused into the GEMM that produces H) [N x 32]
//             A  = [sum_j w_ij U_j, sum_j gx_ij U_j, sum_j gy_ij U_j]   [N x 96]
//             H' = H + SiLU([H, A] Wo^T + bo)                   [N x 128]
//   Y   = H3 Wd^T + bd                                          [N x 16] (8 used)
// with w, gx, gy the dimensionless cubic-spline weight and gradient of each
// edge of the radius-2h graph (h_ij = (h_i + h_j)/2, self excluded),
// computed on the fly from the node positions: w is symmetric and (gx, gy)
// antisymmetric bit for bit.  Dense layers run on MFMA
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation) with the weight tile staged
// in LDS; weights are kept in FP32 (Adam master copy) and in BF16 (W and W^T
// for the backward GEMMs).
#pragma once

#include <stdint.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/mfree_iwf_amd.h"
#include "kernels.h"

namespace mfree {

constexpr int kNF = 16;    // node features
constexpr int kK0 = 32;    // padded input width
constexpr int kH = 128;    // hidden width
constexpr int kC = 32;     // message channels
constexpr int kG = kH + 3 * kC;  // 224: [H, a0, ax, ay]
constexpr int kOut = 8;    // dx, dy, dvx, dvy, dSxx, dSxy, dSyy, drho
constexpr int kOutP = 16;  // padded output width
constexpr int kLayers = 3;
constexpr uint32_t kMaxDeg = 64;  // edge slots per node
constexpr float kSplineNorm = 0.45472840883398667f;  // 10/(7 pi): the 2D cubic spline's factor times h^2

typedef unsigned short bf16_t;  // raw bfloat16 bits

// float -> bfloat16, round to nearest even (finite inputs)
__device__ __host__ inline bf16_t f2bf(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  const uint32_t lsb = (u >> 16) & 1u;
  u += 0x7FFFu + lsb;
  return (bf16_t)(u >> 16);
}
__device__ __host__ inline float bf2f(bf16_t b) {
  const uint32_t u = (uint32_t)b << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

__device__ inline float sigm(float z) { return __builtin_amdgcn_rcpf(1.f + __expf(-z)); }
__device__ inline float silu(float z) { return z * sigm(z); }

// two floats -> packed bf16 pair (low = a), round to nearest even (v_cvt_pk_bf16_f32)
__device__ inline uint32_t pk_bf16(float a, float b) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{a, b}, bf16x2));
}
__device__ inline bf16_t bf16_bits(float a) { return __builtin_bit_cast(bf16_t, (__bf16)a); }
__device__ inline uint2 pack4_bf16(float a, float b, float c, float d) { return make_uint2(pk_bf16(a, b), pk_bf16(c, d)); }

// cubic-spline shape f(q) and f'(q) (src/kernel.cpp:20-55 without the
// 10/(7 pi h^2) factor)
__device__ inline void spline_shape(float q, float& f, float& df) {
  if (q < 1.f) {
    f = 1.f - 1.5f * q * q * (1.f - 0.5f * q);
    df = -3.f * q * (1.f - 0.75f * q);
  } else if (q < 2.f) {
    const float t = 2.f - q;
    f = 0.25f * t * t * t;
    df = -0.75f * t * t;
  } else {
    f = 0.f;
    df = 0.f;
  }
}

// Edge features from node records (x, y, h, -): w = W(r, h_ij) h_ij^2 and
// (gx, gy) = grad_i W(r, h_ij) h_ij^3.  fl(xi - xj) = -fl(xj - xi), so the
// pair (j, i) gets the same w and the negated gradient.  Square root and
// reciprocals are the hardware's single instructions (about 1 ulp).
__device__ inline void edge_weights(const float4 pi, const float4 pj, float& w, float& gx, float& gy) {
  const float dx = pi.x - pj.x, dy = pi.y - pj.y;
  const float hij = 0.5f * (pi.z + pj.z);
  const float r = __builtin_amdgcn_sqrtf(dx * dx + dy * dy);
  float f, df;
  spline_shape(r * __builtin_amdgcn_rcpf(hij), f, df);
  w = kSplineNorm * f;
  const float sc = r > 0.f ? kSplineNorm * df * __builtin_amdgcn_rcpf(r) : 0.f;
  gx = sc * dx;
  gy = sc * dy;
}

// One parameter tensor: FP32 master, Adam moments, gradient, BF16 W and W^T.
struct MpTensor {
  int rows = 0, cols = 0;        // W is rows x cols (out x in)
  size_t off = 0;                // offset into the flat FP32 arrays
  bf16_t* w16 = nullptr;         // rows_p x cols_p (rows to 16, cols to 32, zero padded)
  bf16_t* wt16 = nullptr;        // cols x rows_pk (rows to 32, zero padded)
};

struct MpNet {
  std::vector<MpTensor> t;       // We, be, {Wu, Wo, bo} x L, Wd, bd
  size_t nparam = 0;
  float *w = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;  // flat FP32
  std::vector<void*> allocs;
  int adam_t = 0;
  // tensor indices
  int We, be, Wd, bd;
  int Wu[kLayers], Wo[kLayers], bo[kLayers];
};

// Activations of one forward pass over N nodes (kept for the backward).
// Every array has whole rows of its own (no interleaving of differently
// produced columns), so each kernel writes full cache lines.
struct MpActs {
  uint32_t cap = 0;
  bool train = false;                   // training buffers allocated
  bf16_t* X0 = nullptr;                 // N x 32
  bf16_t* H16[kLayers + 1] = {};        // hidden states, N x 128 bf16
  bf16_t* A16[kLayers] = {};            // aggregated messages of layer l, N x 96 bf16
  bf16_t* U[kLayers] = {};              // N x 32 bf16
  float* Y = nullptr;                   // N x 16
  // training
  float* Z[kLayers + 1] = {};           // pre-activations: Z0 (enc), Z1..Z3 (layers), N x 128
  float* Hf[kLayers + 1] = {};          // hidden states FP32, N x 128
  float* dH = nullptr;                  // N x 128
  bf16_t* dZ16 = nullptr;               // N x 128 (bf16 dZ operand)
  bf16_t* dU16 = nullptr;               // N x 32 (bf16 dY / dU operand)
  bf16_t* dA16 = nullptr;               // N x 96
  float* part = nullptr;                // dW partials
  size_t part_cap = 0;
  std::vector<void*> allocs;
};

// Radius-2h graph over sorted positions, slot-major (ELL): the k-th edge
// of node p is col[k * cap + p], k < deg[p] <= kMaxDeg, so the lanes of a
// wave that handle consecutive nodes read and write consecutive words.
struct MpGraph {
  uint32_t cap = 0;           // nodes (the slot stride)
  uint32_t* deg = nullptr;    // cap
  uint32_t* col = nullptr;    // kMaxDeg * cap (sorted positions)
  float4* npos = nullptr;     // cap: (x, y, h, 0) FP32 by sorted position
  double4* dpos = nullptr;    // cap: (x, y, h, 0) FP64 by sorted position (graph build)
  float* wsum = nullptr;      // cap: sum of w over the edges and the node itself
  std::vector<void*> allocs;
};

mfree_status mp_graph_reserve(MpGraph& g, uint32_t n, std::string& msg);
void mp_graph_free(MpGraph& g);

mfree_status mp_net_init(MpNet& net, uint64_t seed, std::string& msg);
void mp_net_free(MpNet& net);
void mp_net_sync16(MpNet& net, hipStream_t st);  // FP32 master -> BF16 copies
mfree_status mp_acts_reserve(MpActs& a, uint32_t n, bool train, std::string& msg);
void mp_acts_free(MpActs& a);

// Forward over n nodes: features in a.X0, graph g; outputs a.Y.  With
// train == false only the BF16 activations are kept (the residual stream
// is the BF16 copy of H); with train == true the FP32 hidden states and the
// pre-activations the backward needs are stored too.
void mp_forward(const MpNet& net, MpActs& a, const MpGraph& g, uint32_t n, bool train, hipStream_t st);
// The inference forward after the encoder: X0 and U_0 given (a fused
// features + encoder kernel wrote them); the three fused layers and the
// decoder -> a.Y.
void mp_forward_layers(const MpNet& net, MpActs& a, const MpGraph& g, uint32_t n, hipStream_t st);
// Loss + dY from targets T (n x 8, normalised) -> accumulates weight
// gradients into net.g (zeroed by the caller).  loss_out: device float.
void mp_backward(MpNet& net, MpActs& a, const MpGraph& g, uint32_t n, const float* T, float* loss_out,
                 hipStream_t st);
void mp_zero_grad(MpNet& net, hipStream_t st);
void mp_adam(MpNet& net, float lr, float beta1, float beta2, float eps, hipStream_t st);

}  // namespace mfree
