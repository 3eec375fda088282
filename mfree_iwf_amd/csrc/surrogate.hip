// Learned stepper: a message-passing network over the SPH neighbour graph
// that stands in for groups of 4 RK4 steps (no counterpart in the reference,
// whose only integrator is src/rk4.h:71-108).
//
// One learned step of a UL context (every body advances by tau = 4 dt):
//   1. D1 = time derivatives at the state (the stage-0 pass of the UL
//      pipeline, src/body.h:125-148)
//   2. rho and S advanced by D1 over tau (velocity and positions kept); D2 =
//      the acceleration at that state
//   3. base update (FP64):  S += tau*dS1, rho *= exp(tau*drho1/rho),
//                           v += tau*a2, x += tau*xdot1 + tau^2/2*a2
//   4. the network adds a correction to all 8 fields (FP64 update), its
//      inputs being 16 dimensionless node features built from the state,
//      D1, D2 and the velocity one learned step earlier.
// The network is trained on snapshots that the batched solver produces:
// target = (state after 4 RK4 steps) - (base update), made dimensionless
// with h, c0, K and rho0 of the body.
//
// The judged rollout runs groups of learned steps; the last step of each
// group is also computed with 4 plain RK4 steps from the same state.  The
// plain result is kept; if the positions of the two differ by more than the
// threshold (relative L2 over all particles) the group is rolled back and
// redone with plain RK4 steps.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <random>
#include <string>
#include <vector>

#include "ctx.h"
#include "mpnn.h"

using namespace mfree;

namespace mfree {
namespace {

constexpr uint32_t kStepsPerLearned = 4;
constexpr int kNorm = 2 * kNF + 2 * kOut;  // feature mean, std, target mean, std
constexpr float kSplineNorm = 0.45472840883398667f;  // 10/(7 pi): 2D cubic spline factor times h^2

inline uint32_t nblk(uint64_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }

// ---- graph ----------------------------------------------------------------

struct GraphArgs {
  uint32_t n;
  const uint32_t *perm, *key, *scene, *cell_start;
  const SceneConst* sc;
  const SceneGeom* geom;
  const double *x, *y, *h;  // storage order
  uint32_t stride;          // slot stride of col
  uint32_t* deg;
  uint32_t* col;
  float4* npos;
  double4* dpos;
  float* wsum;
  uint32_t* err;
};

// Node records in sorted order: FP64 for the edge predicate (contiguous
// candidate runs), FP32 for the network's edge features.
__global__ __launch_bounds__(256) void k_graph_nodes(GraphArgs a) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.n) return;
  const uint32_t s = a.perm[p];
  const double x = a.x[s], y = a.y[s], h = a.h[s];
  a.dpos[p] = make_double4(x, y, h, 0.0);
  a.npos[p] = make_float4((float)x, (float)y, (float)h, 0.f);
}

// One thread per sorted position p: edges to every other particle of the
// 3x3 cell block with r < 2 h_ij (h_ij = (h_i + h_j)/2), decided in FP64 from
// symmetric differences, so the graph is symmetric.  Writes the slot-major
// edge list, the degree and the kernel sum.
__global__ __launch_bounds__(256) void k_graph(GraphArgs a) {
  const uint32_t p = xcd_tile() * blockDim.x + threadIdx.x;
  if (p >= a.n) return;
  const uint32_t s = a.perm[p];
  const uint32_t b = a.scene[s];
  const SceneGeom g = a.geom[b];
  const uint32_t co = a.sc[b].cell_offset;
  const uint32_t lk = a.key[s] - co;
  const int ix = (int)(lk / g.ny), iy = (int)(lk % g.ny);
  const double4 pi = a.dpos[p];
  const double xi = pi.x, yi = pi.y, hi = pi.z;
  uint32_t d = 0;
  float ws = kSplineNorm;  // the node itself (q = 0)
  // a body whose grid fell back to one cell fails the step: scan no cells
  const int cx0 = max(ix - 1, 0), cx1 = g.err ? cx0 - 1 : min(ix + 1, (int)g.nx - 1);
  const int cy0 = max(iy - 1, 0), cy1 = min(iy + 1, (int)g.ny - 1);
  for (int cx = cx0; cx <= cx1; cx++) {
    const uint32_t c0 = co + (uint32_t)cx * g.ny;
    const uint32_t q0 = a.cell_start[c0 + cy0], q1 = a.cell_start[c0 + cy1 + 1];
    for (uint32_t q = q0; q < q1; q++) {
      if (q == p) continue;
      const double4 pj = a.dpos[q];
      const double dx = xi - pj.x, dy = yi - pj.y;
      // no FMA contraction: the predicate rounds like the FP64 oracle's
      const double r2 = __dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy));
      const double hij = 0.5 * (hi + pj.z);
      if (!(r2 < __dmul_rn(__dmul_rn(4.0, hij), hij))) continue;
      // the weight sum is an FP32 feature: the shape in single precision
      float f, df;
      spline_shape(__builtin_amdgcn_sqrtf((float)r2) * __builtin_amdgcn_rcpf((float)hij), f, df);
      ws += kSplineNorm * f;
      if (d < kMaxDeg) a.col[(size_t)d * a.stride + p] = q;
      d++;
    }
  }
  if (d > kMaxDeg) {
    atomicOr(a.err, EB_NBH);
    d = kMaxDeg;
  }
  a.deg[p] = d;
  a.wsum[p] = ws;
}

// Node records, then (edges) the edge lists; without edges the mixed D1
// pass emits them.
void launch_graph(hipStream_t st, GraphArgs g, MpGraph& mg, bool edges = true) {
  g.stride = mg.cap;
  g.deg = mg.deg;
  g.col = mg.col;
  g.npos = mg.npos;
  g.dpos = mg.dpos;
  g.wsum = mg.wsum;
  hipLaunchKernelGGL(k_graph_nodes, dim3(nblk(g.n, 256)), dim3(256), 0, st, g);
  if (edges) hipLaunchKernelGGL(k_graph, dim3(xcd_grid(nblk(g.n, 256))), dim3(256), 0, st, g);
}

// Edge features of every edge in CSR order (export for checks).
__global__ void k_edge_export(uint32_t n, uint32_t stride, const uint32_t* __restrict__ deg,
                              const uint32_t* __restrict__ col, const uint32_t* __restrict__ rowp,
                              const float4* __restrict__ npos, float* __restrict__ e3) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 pi = npos[i];
  for (uint32_t k = 0; k < deg[i]; k++) {
    float w, gx, gy;
    edge_weights(pi, npos[col[(size_t)k * stride + i]], w, gx, gy);
    const size_t e = rowp[i] + k;
    e3[3 * e] = w;
    e3[3 * e + 1] = gx;
    e3[3 * e + 2] = gy;
  }
}

// ---- learned-step kernels -------------------------------------------------

struct StepArgs {
  uint32_t n;
  const uint32_t *scene, *gid;         // storage order (= graph node order after the base step)
  const SceneConst* sc;
  Fields8 y;                           // state after the base step, storage order
  Fields8 rate;                        // mean rate of the base step: y = y_start + tau rate
  Fields8 k1, k2, k3;                  // its stage rates: RK4 k2, k3 (fields 2..7); Kutta k1 (2, 3), k2 (2..7)
  int order;                           // base step: 3 (Kutta) or 4 (RK4)
  const double *h, *m;                 // storage order
  double* vprev;                       // by global id (2 per particle)
  const float* wsum;                   // by node
  const float* norm;                   // kNorm floats (device), or null
  // outputs
  float* F;                            // raw features by node (kNF), or null
  bf16_t* X0;                          // normalised features by node (kK0), or null
  const float* Y;                      // network output by node (kOutP), or null
  double* base;                        // base update by global id (8), or null
  uint32_t* inv;                       // global id -> node, or null
  double *ox, *oy, *oh;                // start positions / h by node, or null
  const double* cmean;                 // per body: mass-weighted mean of the x, y, vx, vy corrections
  int apply;                           // add the correction to y
  int save_vprev;                      // vprev <- v at the start of the step
};

__device__ inline double tau_of(const SceneConst& c) { return (double)kStepsPerLearned * c.dt; }

// Constants of the base step: every body's dt scaled by kStepsPerLearned.
__global__ void k_step_consts(uint32_t nscene, const SceneConst* __restrict__ sc, SceneConst* __restrict__ out) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nscene) return;
  SceneConst c = sc[s];
  c.dt = tau_of(c);
  c.half_dt = 0.5 * c.dt;
  c.dt6 = c.dt * (1. / 6.);
  out[s] = c;
}

// The 16 raw node features of node p (state field order x, y, vx, vy, rho,
// Sxx, Sxy, Syy).  The base step is an explicit RK step over tau with stage
// increments a_i = tau k_i.  For a linear system y' = A y the stage
// differences give powers of tau A applied to y:
//   Kutta:  D2 = 2 (a2 - a1) = (tau A)^2 y,  D3 = a3 - 2 a2 + a1 = (tau A)^3 y;
//   RK4:    D2 = 2 (a2 - a1),  D3 = 4 (a3 - a2) = (tau A)^3 y,
//           D4 = 4 (a4 - a1 - D2 - D3/2) = 4 (a4 - 2 a3 + a1) = (tau A)^4 y;
// with the mean rate m (k4 = 6 m - k1 - 2 k2 - 2 k3, k3 = 6 m - k1 - 4 k2)
// these need RK4's k2, k3 (D4 = 4 tau (6 m - 2 k2 - 4 k3)) and Kutta's k1 (for
// the velocity's D2) and k2 (D3 = 6 tau (m - k2)).
// the step's error against the exact flow starts with (tau A)^(p+1) y/(p+1)!
// (order p): one more application of the SPH operator to the highest
// difference, which the message passing can learn.  Features:
//   (v - v_prev) / c0 (2), S / K (3), rho/rho0 - 1 (state at the start),
//   Dlo v / c0 (2), Dhi v / c0 (2), Dhi S / K (3), Dhi rho / rho0,
//   c0 tau / h, kernel-weight sum,
// with (Dlo, Dhi) = (D2, D3) for Kutta and (D3, D4) for RK4.
__device__ inline void node_features(const StepArgs& a, uint32_t p, float* f) {
  const uint32_t g = a.gid[p];
  const SceneConst& c = a.sc[a.scene[p]];
  const double tau = tau_of(c);
  const double c0 = sqrt(c.c0sq), K = c.K, rho0 = c.rho0, h = a.h[p];
  double st[8], dlo[8], dhi[8];
#pragma unroll
  for (int k = 2; k < 8; k++) {
    const double m = a.rate.f[k][p], k2 = a.k2.f[k][p];
    st[k] = a.y.f[k][p] - tau * m;
    if (a.order == 4) {
      const double k3 = a.k3.f[k][p];
      dlo[k] = k < 4 ? 4.0 * tau * (k3 - k2) : 0.0;
      dhi[k] = 4.0 * tau * (6.0 * m - 2.0 * k2 - 4.0 * k3);
    } else {
      dlo[k] = k < 4 ? 2.0 * tau * (k2 - a.k1.f[k][p]) : 0.0;
      dhi[k] = 6.0 * tau * (m - k2);
    }
  }
  const double vpx = a.vprev[2 * (size_t)g], vpy = a.vprev[2 * (size_t)g + 1];
  f[0] = (float)((st[2] - vpx) / c0);
  f[1] = (float)((st[3] - vpy) / c0);
  f[2] = (float)(st[5] / K);
  f[3] = (float)(st[6] / K);
  f[4] = (float)(st[7] / K);
  f[5] = (float)(st[4] / rho0 - 1.0);
  f[6] = (float)(dlo[2] / c0);
  f[7] = (float)(dlo[3] / c0);
  f[8] = (float)(dhi[2] / c0);
  f[9] = (float)(dhi[3] / c0);
  f[10] = (float)(dhi[5] / K);
  f[11] = (float)(dhi[6] / K);
  f[12] = (float)(dhi[7] / K);
  f[13] = (float)(dhi[4] / rho0);
  f[14] = (float)(c0 * tau / h);
  f[15] = a.wsum[p];
}

// Network correction of field k (state order) at node p: output (denormalised)
// times the field's scale (h, h, c0, c0, rho0, K, K, K).
__device__ inline double correction(const StepArgs& a, uint32_t p, int k, double h, double c0, double K,
                                    double rho0) {
  const int out_of_field[8] = {0, 1, 2, 3, 7, 4, 5, 6};  // field -> output index
  const double scale[8] = {h, h, c0, c0, rho0, K, K, K};
  const int o = out_of_field[k];
  const double r = (double)a.Y[(size_t)p * kOutP + o] * (double)a.norm[2 * kNF + kOut + o] + (double)a.norm[2 * kNF + o];
  return r * scale[k];
}

// Conservation of the correction.  Kutta's step and the RK4 steps it stands
// in for both conserve each body's linear momentum exactly (antisymmetric
// pair forces) and move its centre of mass by tau times the momentum over
// the mass (the XSPH term sums to zero), so their difference has a zero
// mass-weighted mean in x, y, vx, vy.  The correction is projected onto
// that subspace: per chunk of a body, sums of m dx, m dy, m dvx, m dvy, m
// (fixed-order block reduction), then per body the means.
constexpr int kCorrThreads = 256;
__global__ __launch_bounds__(kCorrThreads) void k_corr_partial(StepArgs a, const Chunk* __restrict__ chunks,
                                                               double* __restrict__ part) {
  const Chunk ch = chunks[blockIdx.x];
  const SceneConst& c = a.sc[ch.scene];
  const double c0 = sqrt(c.c0sq);
  double acc[5] = {0., 0., 0., 0., 0.};
  for (uint32_t p = ch.begin + threadIdx.x; p < ch.end; p += kCorrThreads) {
    const double m = a.m[p], h = a.h[p];
#pragma unroll
    for (int k = 0; k < 4; k++) acc[k] += m * correction(a, p, k, h, c0, c.K, c.rho0);
    acc[4] += m;
  }
  __shared__ double red[5][kCorrThreads];
#pragma unroll
  for (int k = 0; k < 5; k++) red[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int w = kCorrThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
#pragma unroll
      for (int k = 0; k < 5; k++) red[k][threadIdx.x] += red[k][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x < 5) part[5 * (size_t)blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// One workgroup per body: strided sums over its chunks, then the same
// fixed-order tree as k_corr_partial.
__global__ __launch_bounds__(kCorrThreads) void k_corr_final(const uint32_t* __restrict__ chunk_first,
                                                             const double* __restrict__ part,
                                                             double* __restrict__ cmean) {
  const uint32_t s = blockIdx.x;
  double t[5] = {0., 0., 0., 0., 0.};
  for (uint32_t q = chunk_first[s] + threadIdx.x; q < chunk_first[s + 1]; q += kCorrThreads) {
#pragma unroll
    for (int k = 0; k < 5; k++) t[k] += part[5 * (size_t)q + k];
  }
  __shared__ double red[5][kCorrThreads];
#pragma unroll
  for (int k = 0; k < 5; k++) red[k][threadIdx.x] = t[k];
  __syncthreads();
  for (int w = kCorrThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
#pragma unroll
      for (int k = 0; k < 5; k++) red[k][threadIdx.x] += red[k][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x < 4) {
    const double M = red[4][0];
    cmean[4 * (size_t)s + threadIdx.x] = M > 0. ? red[threadIdx.x][0] / M : 0.;
  }
}

// One thread per node: features (training data), base-update export,
// network correction.
__global__ __launch_bounds__(256) void k_learned(StepArgs a) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.n) return;
  const uint32_t g = a.gid[p];
  const SceneConst& c = a.sc[a.scene[p]];
  const double tau = tau_of(c);
  const double c0 = sqrt(c.c0sq), K = c.K, rho0 = c.rho0, h = a.h[p];
  if (a.F) {
    float f[kNF];
    node_features(a, p, f);
#pragma unroll
    for (int k = 0; k < kNF; k++) a.F[(size_t)p * kNF + k] = f[k];
  }
  double nw[8];
#pragma unroll
  for (int k = 0; k < 8; k++) nw[k] = a.y.f[k][p];
  if (a.base) {
#pragma unroll
    for (int k = 0; k < 8; k++) a.base[8 * (size_t)g + k] = nw[k];
  }
  if (a.inv) a.inv[g] = p;
  if (a.ox) {
    a.ox[p] = nw[0] - tau * a.rate.f[0][p];
    a.oy[p] = nw[1] - tau * a.rate.f[1][p];
    a.oh[p] = h;
  }
  if (a.save_vprev) {
    a.vprev[2 * (size_t)g] = nw[2] - tau * a.rate.f[2][p];
    a.vprev[2 * (size_t)g + 1] = nw[3] - tau * a.rate.f[3][p];
  }
  if (a.Y) {
    const double* mean = a.cmean + 4 * (size_t)a.scene[p];
#pragma unroll
    for (int k = 0; k < 8; k++) nw[k] += correction(a, p, k, h, c0, K, rho0) - (k < 4 ? mean[k] : 0.);
  }
  if (a.apply) {
#pragma unroll
    for (int k = 0; k < 8; k++) a.y.f[k][p] = nw[k];
  }
}

// Inference: features and the encoder's messages in one pass.  A wave takes
// 64 nodes (one per lane): it writes their normalised features
// (X0, which the first fused layer reads again) and stages them in LDS, then,
// per 16 nodes, H_0 = SiLU(X0 We^T + be) on v_mfma_f32_16x16x32_bf16
// (D = We X0^T: lane = node lane%16, channels 16t + 4(lane/16)..+3), H_0 rows
// staged as bf16, and U_0 = H_0 Wu_0^T, the only output of the encoder.
constexpr int kFeThreads = 256;
// Grid of the feature kernel: one resident round -- its ~44 KB of LDS per
// workgroup fits 3 workgroups on each of the 256 CUs (measured at 1M nodes:
// 101 us, against 124 us with 1024 workgroups and 106 us with 1536).
constexpr uint32_t kFeBlocksMax = 768;
__global__ __launch_bounds__(kFeThreads) void k_features_encode(StepArgs a, const bf16_t* __restrict__ We,
                                                                const float* __restrict__ be,
                                                                const bf16_t* __restrict__ Wu, bf16_t* __restrict__ U0) {
  // X0 has kNF = 16 features and 16 zero columns: the MFMA's k = 16..31 half
  // is zero, so only the first 16 columns of X0 and We are staged
  constexpr int KE = kNF + 8, MP = kH + 8, KU = kH + 8, T = kH / 16;
  constexpr int kWaves = kFeThreads / 64;
  __shared__ __attribute__((aligned(16))) bf16_t wel[kH * KE];
  __shared__ __attribute__((aligned(16))) bf16_t wul[kC * KU];
  __shared__ __attribute__((aligned(16))) float bel[kH];
  __shared__ __attribute__((aligned(16))) bf16_t xt[kWaves * 64 * KE];
  __shared__ __attribute__((aligned(16))) bf16_t stage[kWaves * 16 * MP];
  for (int i = threadIdx.x; i < kH * 2; i += kFeThreads) {  // 2 chunks of 8 of every We row
    const int r = i >> 1, c = (i & 1) * 8;
    *(uint4*)&wel[r * KE + c] = *(const uint4*)&We[r * kK0 + c];
  }
  for (int i = threadIdx.x * 8; i < kC * kH; i += kFeThreads * 8) {
    const int r = i / kH, c = i - r * kH;
    *(uint4*)&wul[r * KU + c] = *(const uint4*)&Wu[i];
  }
  if (threadIdx.x < kH) bel[threadIdx.x] = be[threadIdx.x];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int q = lane & 15, kq = (lane >> 4) * 8, mq = (lane >> 4) * 4;
  bf16_t* xw = xt + wave * 64 * KE;
  bf16_t* stg = stage + wave * 16 * MP;
  const uint4 zero4 = make_uint4(0u, 0u, 0u, 0u);
  for (uint32_t base = (blockIdx.x * kWaves + wave) * 64u; base < a.n; base += gridDim.x * kWaves * 64u) {
    const uint32_t p = base + lane;
    uint4 row[2] = {zero4, zero4};
    if (p < a.n) {
      float f[kNF];
      node_features(a, p, f);
#pragma unroll
      for (int k = 0; k < kNF; k++) f[k] = (f[k] - a.norm[k]) / a.norm[kNF + k];
#pragma unroll
      for (int k = 0; k < 2; k++)
        row[k] = make_uint4(pk_bf16(f[8 * k], f[8 * k + 1]), pk_bf16(f[8 * k + 2], f[8 * k + 3]),
                            pk_bf16(f[8 * k + 4], f[8 * k + 5]), pk_bf16(f[8 * k + 6], f[8 * k + 7]));
      uint4* xr = (uint4*)&a.X0[(size_t)p * kK0];
      xr[0] = row[0];
      xr[1] = row[1];
      xr[2] = zero4;
      xr[3] = zero4;
    }
    *(uint4*)&xw[lane * KE] = row[0];
    *(uint4*)&xw[lane * KE + 8] = row[1];
    __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
    __builtin_amdgcn_wave_barrier();
#pragma unroll 1
    for (int sub = 0; sub < 4; sub++) {
      // B operand: features kq..kq+7 of node sub*16 + q (zero for kq >= 16)
      const bf16x8 xb = kq < kNF ? *(const bf16x8*)&xw[(sub * 16 + q) * KE + kq] : bf16x8{};
#pragma unroll
      for (int t = 0; t < T; t++) {
        const bf16x8 wa = kq < kNF ? *(const bf16x8*)&wel[(t * 16 + q) * KE + kq] : bf16x8{};
        const f32x4 z = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa, xb, f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
        const int m = t * 16 + mq;
        const float4 bb = *(const float4*)&bel[m];
        *(uint2*)&stg[q * MP + m] = pack4_bf16(silu(z[0] + bb.x), silu(z[1] + bb.y), silu(z[2] + bb.z), silu(z[3] + bb.w));
      }
      __builtin_amdgcn_s_waitcnt(0xc07f);
      __builtin_amdgcn_wave_barrier();
      f32x4 ua[kC / 16];
#pragma unroll
      for (int t2 = 0; t2 < kC / 16; t2++) ua[t2] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kk = 0; kk < kH / 32; kk++) {
        const bf16x8 hb = *(const bf16x8*)&stg[q * MP + kk * 32 + kq];
#pragma unroll
        for (int t2 = 0; t2 < kC / 16; t2++) {
          const bf16x8 wa = *(const bf16x8*)&wul[(t2 * 16 + q) * KU + kk * 32 + kq];
          ua[t2] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa, hb, ua[t2], 0, 0, 0);
        }
      }
      const uint32_t nu = base + sub * 16 + q;
      if (nu < a.n) {
#pragma unroll
        for (int t2 = 0; t2 < kC / 16; t2++)
          *(uint2*)&U0[(size_t)nu * kC + t2 * 16 + mq] = pack4_bf16(ua[t2][0], ua[t2][1], ua[t2][2], ua[t2][3]);
      }
      __builtin_amdgcn_s_waitcnt(0xc07f);
      __builtin_amdgcn_wave_barrier();
    }
  }
}

// Targets of one capture, by the new storage order: (truth - base) / scale,
// outputs ordered dx, dy, dvx, dvy, dSxx, dSxy, dSyy, drho.
__global__ __launch_bounds__(256) void k_targets(uint32_t n, Fields8 y, const uint32_t* __restrict__ gid,
                                                 const uint32_t* __restrict__ scene, const SceneConst* __restrict__ sc,
                                                 const double* __restrict__ h, const double* __restrict__ base,
                                                 const uint32_t* __restrict__ inv, float* __restrict__ T) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t g = gid[i];
  const SceneConst& c = sc[scene[i]];
  const double c0 = sqrt(c.c0sq);
  const double scale[8] = {h[i], h[i], c0, c0, c.rho0, c.K, c.K, c.K};
  const int out_of_field[8] = {0, 1, 2, 3, 7, 4, 5, 6};
  float* t = T + (size_t)inv[g] * kOut;
#pragma unroll
  for (int k = 0; k < 8; k++) t[out_of_field[k]] = (float)((y.f[k][i] - base[8 * (size_t)g + k]) / scale[k]);
}

__global__ void k_record_vprev(uint32_t n, const double* __restrict__ vx, const double* __restrict__ vy,
                               const uint32_t* __restrict__ gid, double* __restrict__ vprev) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t g = gid[i];
  vprev[2 * (size_t)g] = vx[i];
  vprev[2 * (size_t)g + 1] = vy[i];
}

// Column sums and sums of squares of F (kNF) and T (kOut) over `rows`
// rows, accumulated in FP64 into acc[2 * (kNF + kOut)].
constexpr uint32_t kStatBlocks = 128;
__global__ __launch_bounds__(256) void k_stats(uint64_t rows, const float* __restrict__ F,
                                               const float* __restrict__ T, double* __restrict__ acc) {
  constexpr int C = kNF + kOut;
  __shared__ double sh[256 / 64][2 * C];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double s1[C], s2[C];
#pragma unroll
  for (int k = 0; k < C; k++) s1[k] = s2[k] = 0.0;
  for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < rows; r += (uint64_t)gridDim.x * blockDim.x) {
#pragma unroll
    for (int k = 0; k < kNF; k++) {
      const double v = F[r * kNF + k];
      s1[k] += v;
      s2[k] += v * v;
    }
#pragma unroll
    for (int k = 0; k < kOut; k++) {
      const double v = T[r * kOut + k];
      s1[kNF + k] += v;
      s2[kNF + k] += v * v;
    }
  }
#pragma unroll
  for (int k = 0; k < C; k++) {
    for (int o = 32; o > 0; o >>= 1) {
      s1[k] += __shfl_xor(s1[k], o, 64);
      s2[k] += __shfl_xor(s2[k], o, 64);
    }
  }
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < C; k++) {
      sh[w][k] = s1[k];
      sh[w][C + k] = s2[k];
    }
  }
  __syncthreads();
  if (threadIdx.x < 2 * C) {
    double t = 0.0;
    for (int k = 0; k < 256 / 64; k++) t += sh[k][threadIdx.x];
    atomicAdd(&acc[threadIdx.x], t);
  }
}

// Largest |rho/rho0 - 1| per body (float bits, atomic max; NaN compares
// above every finite value).
__global__ void k_body_compression(uint32_t n, const double* __restrict__ rho, const uint32_t* __restrict__ scene,
                                   const SceneConst* __restrict__ sc, uint32_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t b = scene[i];
  const float d = (float)fabs(rho[i] / sc[b].rho0 - 1.0);
  uint32_t u;
  memcpy(&u, &d, 4);
  atomicMax(&out[b], u);
}

// Training minibatch inputs: X0[p] = bf16((F[r] - mean) / std), Tn[p] =
// (T[r] - tmean) / tstd with r = perm[p].
__global__ __launch_bounds__(256) void k_mb_inputs(uint32_t n, const uint32_t* __restrict__ perm,
                                                   const float* __restrict__ F, const float* __restrict__ T,
                                                   const float* __restrict__ norm, bf16_t* __restrict__ X0,
                                                   float* __restrict__ Tn) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint32_t r = perm[p];
  bf16_t* x = X0 + (size_t)p * kK0;
#pragma unroll
  for (int k = 0; k < kNF; k++) x[k] = f2bf((F[(size_t)r * kNF + k] - norm[k]) / norm[kNF + k]);
#pragma unroll
  for (int k = kNF; k < kK0; k++) x[k] = 0;
#pragma unroll
  for (int k = 0; k < kOut; k++)
    Tn[(size_t)p * kOut + k] = (T[(size_t)r * kOut + k] - norm[2 * kNF + k]) / norm[2 * kNF + kOut + k];
}

// scene[i] for the bodies of a training window (offsets in sc[].offset).
__global__ void k_scene_ids(uint32_t n, const SceneConst* __restrict__ sc, uint32_t ns, uint32_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t lo = 0, hi = ns;  // last s with sc[s].offset <= i
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (sc[mid].offset <= i)
      lo = mid;
    else
      hi = mid;
  }
  out[i] = lo;
}

__global__ void k_iota(uint32_t n, uint32_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = i;
}

// Judge: the 8 state fields of storage buffer y by global id (out: 8 x N,
// field-major).
__global__ void k_state_by_gid(uint32_t n, Fields8 y, const uint32_t* __restrict__ gid, double* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t g = gid[i];
#pragma unroll
  for (int k = 0; k < 8; k++) out[k * (size_t)n + g] = y.f[k][i];
}

// Judge: per chunk of one body (Chunk ranges of the storage buffer y),
// partial sums for the field groups positions (x, y), velocities, density
// and stresses of: the squared difference of the learned and the plain
// increments, the same for the base step's increment (base, 8 per gid), the
// squared plain increment, and the squared judged state.  Increments are
// taken from the judged state s0 (8 x N by gid); cand is the learned step's
// end state (8 x N by gid); y is the plain result.
constexpr int kJudgeGroups = 4;
constexpr int kJudgeSums = 4;
// A judged step whose network correction leaves the velocities, density or
// stresses of a body further from the plain steps than this many times the
// base step alone is rejected.
constexpr double kNetRatioLimit = 2.0;
constexpr double kNetFloor = 1e-5;
__device__ inline int judge_group(int k) { return k < 2 ? 0 : (k < 4 ? 1 : (k == 4 ? 2 : 3)); }
__global__ __launch_bounds__(256) void k_judge_partial(const Chunk* __restrict__ ch, uint32_t n, Fields8 y,
                                                       const uint32_t* __restrict__ gid, const double* __restrict__ s0,
                                                       const double* __restrict__ cand,
                                                       const double* __restrict__ base, double* __restrict__ part) {
  constexpr int kS = kJudgeSums * kJudgeGroups;
  const Chunk c = ch[blockIdx.x];
  double acc[kS];
#pragma unroll
  for (int k = 0; k < kS; k++) acc[k] = 0.0;
  for (uint32_t i = c.begin + threadIdx.x; i < c.end; i += blockDim.x) {
    const size_t g = gid[i];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const double a = s0[k * (size_t)n + g];
      const double dp = y.f[k][i] - a, dl = cand[k * (size_t)n + g] - a, db = base[8 * g + k] - a;
      const int q = judge_group(k);
      acc[kJudgeSums * q] += (dl - dp) * (dl - dp);
      acc[kJudgeSums * q + 1] += (db - dp) * (db - dp);
      acc[kJudgeSums * q + 2] += dp * dp;
      acc[kJudgeSums * q + 3] += a * a;
    }
  }
  __shared__ double red[kS][256];
#pragma unroll
  for (int k = 0; k < kS; k++) red[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (uint32_t w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int k = 0; k < kS; k++) red[k][threadIdx.x] += red[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < kS) part[blockIdx.x * kS + threadIdx.x] = red[threadIdx.x][0];
}

}  // namespace
}  // namespace mfree

// ---------------------------------------------------------------------------

struct mfree_surrogate {
  mfree_surrogate_options opt;
  std::string err;
  hipStream_t st = nullptr;
  std::mt19937_64 rng;
  MpNet net;
  MpActs acts;
  MpGraph graph;
  float* dnorm = nullptr;  // kNorm
  float hnorm[kNorm];
  bool have_norm = false;

  // dataset (rows = nodes of all captures, each capture in its sorted order)
  uint64_t rows = 0, rows_cap = 0;
  float *dsF = nullptr, *dsT = nullptr;
  double *dsX = nullptr, *dsY = nullptr, *dsH = nullptr;
  struct Capture {
    uint64_t row0;
    std::vector<uint32_t> off;  // body offsets within the capture (S+1)
    std::vector<uint8_t> valid; // body usable for training (S)
  };
  std::vector<Capture> caps;
  // training windows: consecutive bodies of one capture, <= max_batch_nodes
  struct Window {
    uint64_t row0;
    uint32_t n, ns, nch, sc_off, ch_off, cf_off;
  };
  std::vector<Window> wins;
  bool wins_ready = false;
  std::vector<void*> wallocs;  // window tables
  SceneConst* wsc = nullptr;
  Chunk* wch = nullptr;
  uint32_t* wcf = nullptr;
  uint32_t tcap = 0, tscene_cap = 0, tchunk_cap = 0;
  std::vector<void*> tallocs;  // training binning buffers
  uint32_t *tkey = nullptr, *trank = nullptr, *tcount = nullptr, *tcell = nullptr, *tperm = nullptr,
           *ttsum = nullptr, *tscene = nullptr, *tiota = nullptr, *terr = nullptr;
  SceneGeom* tgeom = nullptr;
  double* tpart = nullptr;
  float* tT = nullptr;
  float* loss = nullptr;
  uint32_t loss_cap = 0;

  // per-context scratch
  uint64_t bound = 0;  // serial of the bound context (0: none)
  uint32_t bN = 0;
  uint64_t bver = 0;   // state version of the bound context when uniform_h was determined
  bool uniform_h = false;  // every body of the bound context has one smoothing length
  std::vector<void*> callocs;
  double *base = nullptr, *vprev = nullptr;
  double *jstart = nullptr, *cand = nullptr, *devp = nullptr;  // judge: judged state, learned end (8 x N by gid),
                                                               // per-chunk sums
  SceneConst* sc4 = nullptr;  // constants of the base step (dt scaled by kStepsPerLearned)
  Fields8 k1s{}, k2s{}, k3s{};  // stage rates of the base step (storage order; fields 2..7)
  double *cpart = nullptr, *cmean = nullptr;  // correction sums per chunk (5), means per body (4)
  uint32_t* inv = nullptr;
  uint32_t* comp = nullptr;  // per body: largest compression (float bits)
  bool vprev_valid = false;
  // rollback snapshot (storage buffer of the group start)
  double* snap = nullptr;   // 10 fields x N: y(8), m, h
  uint32_t* snapu = nullptr;  // scene, gid
  SceneConst* snap_sc = nullptr;
  SceneConst* jsc = nullptr;
  hipGraphExec_t lexec[2] = {nullptr, nullptr};  // captured learned steps per starting storage buffer
};

namespace mfree {
namespace {

mfree_status sfail(mfree_surrogate* s, mfree_status st, const std::string& m) {
  if (s) s->err = m;
  return st;
}

#define SCHK(s, call)                                                                        \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess) return sfail(s, MFREE_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
hipError_t salloc(std::vector<void*>& list, T** p, size_t count) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
  if (e == hipSuccess) {
    list.push_back(q);
    *p = (T*)q;
  }
  return e;
}

void free_list(std::vector<void*>& l) {
  for (void* p : l) (void)hipFree(p);
  l.clear();
}

void drop_graphs(mfree_surrogate* s) {
  for (int b = 0; b < 2; b++)
    if (s->lexec[b]) {
      (void)hipGraphExecDestroy(s->lexec[b]);
      s->lexec[b] = nullptr;
    }
}

// Whether every body of the context has one smoothing length (h changes only
// through mfree_set_state, which advances the context's state version).
mfree_status check_uniform_h(mfree_surrogate* s, mfree_ctx* c) {
  const size_t N = c->N;
  std::vector<double> hh(N);
  SCHK(s, hipMemcpy(hh.data(), c->h[c->cur], N * sizeof(double), hipMemcpyDeviceToHost));
  bool u = true;
  for (size_t b = 0; b < c->S && u; b++)
    for (uint32_t k = c->off[b] + 1; k < c->off[b + 1] && u; k++) u = hh[k] == hh[c->off[b]];
  s->uniform_h = u;
  return MFREE_OK;
}

// Per-context buffers (base update, previous velocities, snapshots); reallocated when the surrogate meets another context.
mfree_status bind_ctx(mfree_surrogate* s, mfree_ctx* c) {
  mfree_status st = commit(c);
  if (st != MFREE_OK) return sfail(s, st, std::string("context: ") + c->err);
  if (c->opt.method != MFREE_METHOD_UL_MONAGHAN)
    return sfail(s, MFREE_ERR_UNSUPPORTED, "the learned stepper drives UL (Gray-Monaghan) contexts");
  if (c->opt.isolate_bodies)
    return sfail(s, MFREE_ERR_UNSUPPORTED, "the learned stepper's rollback covers the whole context: no isolated bodies");
  if (c->opt.device != s->opt.device) return sfail(s, MFREE_ERR_ARG, "context and surrogate use different devices");
  if (s->bound == c->serial && s->bN == c->N) {
    if (s->bver == c->state_version) return MFREE_OK;
    // new state written into the context: h may have changed, and with it
    // the pair pass the base step may emit the graph from
    SCHK(s, hipStreamSynchronize(c->stream));
    drop_graphs(s);
    mfree_status st2 = check_uniform_h(s, c);
    if (st2 != MFREE_OK) return st2;
    s->bver = c->state_version;
    return MFREE_OK;
  }
  SCHK(s, hipDeviceSynchronize());
  drop_graphs(s);
  free_list(s->callocs);
  const size_t N = c->N;
  SCHK(s, salloc(s->callocs, &s->base, 8 * N));
  SCHK(s, salloc(s->callocs, &s->vprev, 2 * N));
  SCHK(s, salloc(s->callocs, &s->jstart, 8 * N));
  SCHK(s, salloc(s->callocs, &s->cand, 8 * N));
  SCHK(s, salloc(s->callocs, &s->devp, kJudgeSums * kJudgeGroups * (size_t)c->nchunks));
  SCHK(s, salloc(s->callocs, &s->inv, N));
  SCHK(s, salloc(s->callocs, &s->comp, c->S));
  SCHK(s, salloc(s->callocs, &s->snap, 10 * N));
  SCHK(s, salloc(s->callocs, &s->snapu, 2 * N));
  SCHK(s, salloc(s->callocs, &s->sc4, c->S));
  SCHK(s, salloc(s->callocs, &s->cpart, 5 * (size_t)c->nchunks));
  SCHK(s, salloc(s->callocs, &s->cmean, 4 * (size_t)c->S));
  // the stage rates the features use (node_features): RK4 k2, k3 of v, rho,
  // S; Kutta k1 of v and k2 of v, rho, S
  s->k1s = s->k2s = s->k3s = Fields8{};
  for (int k = 2; k < 8; k++) {
    SCHK(s, salloc(s->callocs, &s->k2s.f[k], N));
    if (s->opt.base_order == 4) SCHK(s, salloc(s->callocs, &s->k3s.f[k], N));
    else if (k < 4) SCHK(s, salloc(s->callocs, &s->k1s.f[k], N));
  }
  SCHK(s, salloc(s->callocs, &s->snap_sc, c->S));
  SCHK(s, salloc(s->callocs, &s->jsc, c->S));
  std::string msg;
  if (mp_graph_reserve(s->graph, c->N, msg) != MFREE_OK || mp_acts_reserve(s->acts, c->N, false, msg) != MFREE_OK)
    return sfail(s, MFREE_ERR_HIP, msg);
  st = check_uniform_h(s, c);
  if (st != MFREE_OK) return st;
  s->bound = c->serial;
  s->bN = c->N;
  s->bver = c->state_version;
  s->vprev_valid = false;
  return MFREE_OK;
}

// Learned-step arguments over the current storage buffer (the one the base
// step wrote: storage index = graph node).
StepArgs step_args(mfree_surrogate* s, mfree_ctx* c) {
  StepArgs a;
  memset(&a, 0, sizeof(a));
  const int b = c->cur;
  a.n = c->N;
  a.scene = c->scene[b];
  a.gid = c->gid[b];
  a.sc = c->sc;
  a.y = c->y[b];
  a.rate = c->acc;
  a.k1 = s->k1s;
  a.k2 = s->k2s;
  a.k3 = s->k3s;
  a.order = s->opt.base_order;
  a.h = c->h[b];
  a.m = c->m[b];
  a.vprev = s->vprev;
  a.wsum = s->graph.wsum;
  a.norm = s->dnorm;
  return a;
}

void record_vprev(mfree_surrogate* s, mfree_ctx* c) {
  const int b = c->cur;
  hipLaunchKernelGGL(k_record_vprev, dim3(nblk(c->N, 256)), dim3(256), 0, c->stream, c->N, c->y[b].f[2],
                     c->y[b].f[3], c->gid[b], s->vprev);
  s->vprev_valid = true;
}

// Graph of the state in storage buffer b over the current cell list: node
// records, and the edge lists unless a pair pass emits them.
void enqueue_graph(mfree_surrogate* s, mfree_ctx* c, int b, bool edges) {
  GraphArgs g;
  memset(&g, 0, sizeof(g));
  g.n = c->N;
  g.perm = c->perm;
  g.key = c->key;
  g.scene = c->scene[b];
  g.cell_start = c->cell_start;
  g.sc = c->sc;
  g.geom = c->geom;
  g.x = c->y[b].f[0];
  g.y = c->y[b].f[1];
  g.h = c->h[b];
  g.err = c->errb;
  launch_graph(c->stream, g, s->graph, edges);
}

// Base step of a learned step: one RK step (RK4, or Kutta's third-order
// method with base_order 3) over tau = kStepsPerLearned dt that leaves the
// mean rate in acc and the stage rates k1, k2 (and k3 of RK4) in k1s, k2s,
// k3s.  graph: the network's graph of the start state over the stage-1 cell
// list (nodes = stage-1 sorted positions = storage indices of
// the new buffer); on a mixed context whose bodies each have one h, the
// stage-1 pair pass emits the edge lists.
void enqueue_base(mfree_surrogate* s, mfree_ctx* c, bool graph, bool advance_time) {
  hipLaunchKernelGGL(k_step_consts, dim3((c->S + 63) / 64), dim3(64), 0, c->stream, c->S, c->sc, s->sc4);
  RkPlan plan;
  plan.sc = s->sc4;
  plan.order = s->opt.base_order;
  plan.rate_out = true;
  plan.k1s = s->k1s;
  plan.k2s = s->k2s;
  plan.k3s = s->k3s;
  plan.advance_time = false;
  if (graph) {
    const bool fused = c->opt.precision == MFREE_PRECISION_MIXED && s->uniform_h;
    plan.after_bin1 = [s, c, fused](int b) { enqueue_graph(s, c, b, !fused); };
    if (fused)
      plan.stage1 = [s](UlArgs& u) {
        u.gcol = s->graph.col;
        u.gdeg = s->graph.deg;
        u.gwsum = s->graph.wsum;
        u.gdpos = s->graph.dpos;
        u.gstride = s->graph.cap;
      };
  }
  enqueue_ul_rk(c, plan, nullptr);
  if (advance_time)
    hipLaunchKernelGGL(k_advance_time_n, dim3((c->S + 63) / 64), dim3(64), 0, c->stream, c->S, c->sc,
                       kStepsPerLearned);
}

void advance_time_host(mfree_ctx* c, uint32_t nsteps) {
  for (uint32_t b = 0; b < c->S; b++)
    for (uint32_t k = 0; k < nsteps; k++) c->time[b] = c->time[b] + c->prm[b].dt;
}

// Device work of one learned step of every body (capturable: kernels only);
// storage moves to the other buffer.
void enqueue_learned_device(mfree_surrogate* s, mfree_ctx* c, bool use_net, double* keep_base = nullptr) {
  enqueue_base(s, c, use_net, true);
  StepArgs a = step_args(s, c);
  a.base = keep_base;  // the base step's result by gid, before the correction
  if (use_net) {
    a.X0 = s->acts.X0;
    // features and the encoder first (they read the state), then the fused
    // layers, then the correction
    const MpNet& net = s->net;
    const uint32_t fe_blocks = std::max<uint32_t>(1, std::min<uint32_t>(nblk(c->N, kFeThreads), kFeBlocksMax));
    hipLaunchKernelGGL(k_features_encode, dim3(fe_blocks), dim3(kFeThreads), 0, c->stream, a,
                       net.t[net.We].w16, net.w + net.t[net.be].off, net.t[net.Wu[0]].w16, s->acts.U[0]);
    mp_forward_layers(s->net, s->acts, s->graph, c->N, c->stream);
    a.X0 = nullptr;
    a.Y = s->acts.Y;
    a.cmean = s->cmean;
    hipLaunchKernelGGL(k_corr_partial, dim3(c->nchunks), dim3(kCorrThreads), 0, c->stream, a, c->chunks, s->cpart);
    hipLaunchKernelGGL(k_corr_final, dim3(c->S), dim3(kCorrThreads), 0, c->stream, c->chunk_first, s->cpart,
                       s->cmean);
  }
  a.apply = use_net ? 1 : 0;
  a.save_vprev = 1;
  hipLaunchKernelGGL(k_learned, dim3(nblk(c->N, 256)), dim3(256), 0, c->stream, a);
}

// One learned step of every body (no synchronisation).  With graphs enabled
// on the context, the network step is replayed from a graph captured once
// per starting storage buffer.
mfree_status enqueue_learned(mfree_surrogate* s, mfree_ctx* c, bool use_net, double* keep_base = nullptr) {
  const int a = c->cur;
  if (use_net && c->opt.use_graph && !keep_base) {
    if (!s->lexec[a]) {
      hipGraph_t g;
      SCHK(s, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      enqueue_learned_device(s, c, true);
      c->cur = a;  // capture did not execute anything
      SCHK(s, hipStreamEndCapture(c->stream, &g));
      hipError_t e = hipGraphInstantiate(&s->lexec[a], g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) {
        s->lexec[a] = nullptr;
        return sfail(s, MFREE_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
      }
    }
    SCHK(s, hipGraphLaunch(s->lexec[a], c->stream));
    c->cur = 1 - a;
  } else {
    enqueue_learned_device(s, c, use_net, keep_base);
  }
  advance_time_host(c, kStepsPerLearned);
  return MFREE_OK;
}

mfree_status ensure_rows(mfree_surrogate* s, uint64_t need) {
  if (need <= s->rows_cap) return MFREE_OK;
  SCHK(s, hipDeviceSynchronize());
  float *F = nullptr, *T = nullptr;
  double *X = nullptr, *Y = nullptr, *H = nullptr;
  SCHK(s, hipMalloc(&F, need * kNF * sizeof(float)));
  SCHK(s, hipMalloc(&T, need * kOut * sizeof(float)));
  SCHK(s, hipMalloc(&X, need * sizeof(double)));
  SCHK(s, hipMalloc(&Y, need * sizeof(double)));
  SCHK(s, hipMalloc(&H, need * sizeof(double)));
  if (s->rows) {
    SCHK(s, hipMemcpy(F, s->dsF, s->rows * kNF * sizeof(float), hipMemcpyDeviceToDevice));
    SCHK(s, hipMemcpy(T, s->dsT, s->rows * kOut * sizeof(float), hipMemcpyDeviceToDevice));
    SCHK(s, hipMemcpy(X, s->dsX, s->rows * sizeof(double), hipMemcpyDeviceToDevice));
    SCHK(s, hipMemcpy(Y, s->dsY, s->rows * sizeof(double), hipMemcpyDeviceToDevice));
    SCHK(s, hipMemcpy(H, s->dsH, s->rows * sizeof(double), hipMemcpyDeviceToDevice));
  }
  (void)hipFree(s->dsF);
  (void)hipFree(s->dsT);
  (void)hipFree(s->dsX);
  (void)hipFree(s->dsY);
  (void)hipFree(s->dsH);
  s->dsF = F;
  s->dsT = T;
  s->dsX = X;
  s->dsY = Y;
  s->dsH = H;
  s->rows_cap = need;
  return MFREE_OK;
}

mfree_status build_windows(mfree_surrogate* s);

// Statistics of the training rows (the windows) -> normalisation (mean,
// std per column).
mfree_status compute_norm(mfree_surrogate* s) {
  constexpr int C = kNF + kOut;
  mfree_status st = build_windows(s);
  if (st != MFREE_OK) return st;
  double* acc = nullptr;
  SCHK(s, hipMalloc(&acc, 2 * C * sizeof(double)));
  hipError_t e = hipMemsetAsync(acc, 0, 2 * C * sizeof(double), s->st);
  uint64_t rows = 0;
  for (const auto& w : s->wins) {
    hipLaunchKernelGGL(k_stats, dim3(kStatBlocks), dim3(256), 0, s->st, (uint64_t)w.n, s->dsF + w.row0 * kNF,
                       s->dsT + w.row0 * kOut, acc);
    rows += w.n;
  }
  double h[2 * C];
  if (e == hipSuccess) e = hipMemcpyAsync(h, acc, sizeof(h), hipMemcpyDeviceToHost, s->st);
  if (e == hipSuccess) e = hipStreamSynchronize(s->st);
  (void)hipFree(acc);
  if (e != hipSuccess) return sfail(s, MFREE_ERR_HIP, hipGetErrorString(e));
  const double n = (double)rows;
  for (int k = 0; k < C; k++) {
    const double mean = h[k] / n;
    const double var = std::max(h[C + k] / n - mean * mean, 0.0);
    double sd = sqrt(var);
    // a column that is constant up to rounding keeps unit scale
    if (!(sd > 1e-6 * fabs(mean) + 1e-30) || !isfinite(sd)) sd = std::max(fabs(mean), 1.0);
    if (k < kNF) {
      s->hnorm[k] = (float)mean;
      s->hnorm[kNF + k] = (float)sd;
    } else {
      s->hnorm[2 * kNF + (k - kNF)] = (float)mean;
      s->hnorm[2 * kNF + kOut + (k - kNF)] = (float)sd;
    }
  }
  SCHK(s, hipMemcpy(s->dnorm, s->hnorm, sizeof(s->hnorm), hipMemcpyHostToDevice));
  s->have_norm = true;
  return MFREE_OK;
}

// Training windows over the captures and the binning buffers they need.
mfree_status build_windows(mfree_surrogate* s) {
  if (s->wins_ready) return MFREE_OK;
  SCHK(s, hipDeviceSynchronize());
  free_list(s->wallocs);
  s->wins.clear();
  std::vector<SceneConst> hsc;
  std::vector<Chunk> hch;
  std::vector<uint32_t> hcf;
  const uint32_t cap = std::max<uint32_t>(s->opt.max_batch_nodes, 1);
  uint32_t max_n = 0, max_ns = 0, max_ch = 0;
  for (const auto& cp : s->caps) {
    const uint32_t S = (uint32_t)cp.off.size() - 1;
    uint32_t b = 0;
    while (b < S) {
      if (!cp.valid[b]) {
        b++;
        continue;
      }
      uint32_t e = b + 1;
      while (e < S && cp.valid[e] && cp.off[e + 1] - cp.off[b] <= cap) e++;
      mfree_surrogate::Window w;
      w.row0 = cp.row0 + cp.off[b];
      w.n = cp.off[e] - cp.off[b];
      w.ns = e - b;
      w.sc_off = (uint32_t)hsc.size();
      w.ch_off = (uint32_t)hch.size();
      w.cf_off = (uint32_t)hcf.size();
      for (uint32_t k = b; k < e; k++) {
        SceneConst c;
        memset(&c, 0, sizeof(c));
        c.offset = cp.off[k] - cp.off[b];
        c.n = cp.off[k + 1] - cp.off[k];
        c.cell_offset = c.offset;
        hsc.push_back(c);
        hcf.push_back((uint32_t)(hch.size() - w.ch_off));
        for (uint32_t i = c.offset; i < c.offset + c.n; i += kChunk)
          hch.push_back(Chunk{k - b, i, std::min(i + kChunk, c.offset + c.n), 0});
      }
      hcf.push_back((uint32_t)(hch.size() - w.ch_off));
      w.nch = (uint32_t)(hch.size() - w.ch_off);
      s->wins.push_back(w);
      max_n = std::max(max_n, w.n);
      max_ns = std::max(max_ns, w.ns);
      max_ch = std::max(max_ch, w.nch);
      b = e;
    }
  }
  if (s->wins.empty()) return sfail(s, MFREE_ERR_ARG, "no usable training snapshots (collect first)");
  SCHK(s, salloc(s->wallocs, &s->wsc, hsc.size()));
  SCHK(s, salloc(s->wallocs, &s->wch, hch.size()));
  SCHK(s, salloc(s->wallocs, &s->wcf, hcf.size()));
  SCHK(s, hipMemcpy(s->wsc, hsc.data(), hsc.size() * sizeof(SceneConst), hipMemcpyHostToDevice));
  SCHK(s, hipMemcpy(s->wch, hch.data(), hch.size() * sizeof(Chunk), hipMemcpyHostToDevice));
  SCHK(s, hipMemcpy(s->wcf, hcf.data(), hcf.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  if (max_n > s->tcap || max_ns > s->tscene_cap || max_ch > s->tchunk_cap) {
    free_list(s->tallocs);
    const size_t n = max_n;
    SCHK(s, salloc(s->tallocs, &s->tkey, n));
    SCHK(s, salloc(s->tallocs, &s->trank, n));
    SCHK(s, salloc(s->tallocs, &s->tcount, n));
    SCHK(s, hipMemset(s->tcount, 0, n * sizeof(uint32_t)));  // the binning keeps it zero between calls
    SCHK(s, salloc(s->tallocs, &s->tcell, n + 1));
    SCHK(s, salloc(s->tallocs, &s->tperm, n));
    SCHK(s, salloc(s->tallocs, &s->ttsum, (n + kScanTile - 1) / kScanTile + 1));
    SCHK(s, salloc(s->tallocs, &s->tscene, n));
    SCHK(s, salloc(s->tallocs, &s->tiota, n));
    SCHK(s, salloc(s->tallocs, &s->terr, 1));
    SCHK(s, salloc(s->tallocs, &s->tgeom, max_ns));
    SCHK(s, salloc(s->tallocs, &s->tpart, 5 * (size_t)max_ch));
    SCHK(s, salloc(s->tallocs, &s->tT, n * kOut));
    SCHK(s, hipMemset(s->terr, 0, sizeof(uint32_t)));
    hipLaunchKernelGGL(k_iota, dim3(nblk(n, 256)), dim3(256), 0, s->st, (uint32_t)n, s->tiota);
    s->tcap = max_n;
    s->tscene_cap = max_ns;
    s->tchunk_cap = max_ch;
  }
  std::string msg;
  drop_graphs(s);  // the reservations below may move the buffers a captured step uses
  if (mp_graph_reserve(s->graph, max_n, msg) != MFREE_OK || mp_acts_reserve(s->acts, max_n, true, msg) != MFREE_OK)
    return sfail(s, MFREE_ERR_HIP, msg);
  s->wins_ready = true;
  return MFREE_OK;
}

// One Adam step on window w; the loss of the minibatch is added to *loss.
void enqueue_train_step(mfree_surrogate* s, const mfree_surrogate::Window& w, float lr, float* loss) {
  hipStream_t st = s->st;
  BinArgs a;
  memset(&a, 0, sizeof(a));
  a.n = w.n;
  a.nscene = w.ns;
  a.nchunks = w.nch;
  a.ncell_cap = w.n;
  a.deterministic = 1;
  a.kernel_width = 2.0;
  a.cell_min = 0.0;
  a.chunks = s->wch + w.ch_off;
  a.chunk_first = s->wcf + w.cf_off;
  a.sc = s->wsc + w.sc_off;
  a.geom = s->tgeom;
  a.x = s->dsX + w.row0;
  a.y = s->dsY + w.row0;
  a.h = s->dsH + w.row0;
  a.scene = s->tscene;
  a.gid = s->tiota;
  a.part = s->tpart;
  a.key = s->tkey;
  a.rank = s->trank;
  a.count = s->tcount;
  a.tsum = s->ttsum;
  a.cell_start = s->tcell;
  a.perm = s->tperm;
  a.err = s->terr;
  hipLaunchKernelGGL(k_scene_ids, dim3(nblk(w.n, 256)), dim3(256), 0, st, w.n, a.sc, w.ns, s->tscene);
  launch_binning(st, a);
  GraphArgs g;
  memset(&g, 0, sizeof(g));
  g.n = w.n;
  g.perm = s->tperm;
  g.key = s->tkey;
  g.scene = s->tscene;
  g.cell_start = s->tcell;
  g.sc = a.sc;
  g.geom = s->tgeom;
  g.x = a.x;
  g.y = a.y;
  g.h = a.h;
  g.err = s->terr;
  launch_graph(st, g, s->graph);
  hipLaunchKernelGGL(k_mb_inputs, dim3(nblk(w.n, 256)), dim3(256), 0, st, w.n, s->tperm, s->dsF + w.row0 * kNF,
                     s->dsT + w.row0 * kOut, s->dnorm, s->acts.X0, s->tT);
  mp_forward(s->net, s->acts, s->graph, w.n, true, st);
  mp_zero_grad(s->net, st);
  mp_backward(s->net, s->acts, s->graph, w.n, s->tT, loss, st);
  mp_adam(s->net, lr, 0.9f, 0.999f, 1e-8f, st);
}

mfree_status sync_ctx(mfree_surrogate* s, mfree_ctx* c) {
  mfree_status st = check_err_bits(c);
  if (st != MFREE_OK) return sfail(s, st, std::string("learned step: ") + c->err);
  return MFREE_OK;
}

// Judge of a learned step, per body, from the increments over the judged
// interval of the learned step (u_L), of its own base step (u_B, the same
// step without the network's correction) and of the plain RK4 steps (u_P),
// all from the judged state u_0:
//   positions: || (x_L - x_0) - (x_P - x_0) || / || x_P - x_0 ||  (floored at
//              1e-6 || x_0 ||): the relative error of the displacement;
//   velocities, density, stresses: || u_L - u_P || / || u_B - u_P ||, how
//              many times the network's correction leaves the step further
//              from the plain steps than the base step alone is (the
//              network should bring it closer: a ratio below 1).  Where the
//              base step is nearly exact (a body in free flight), the
//              denominator is floored at kNetFloor of the body's scale of the
//              field (c0, rho0, K: the units of the network's output) per
//              particle, so that errors below that level always pass.
// out[0..3]: the largest of each measure over the bodies; *worst_body: the
// body of the largest position deviation or ratio excess.
struct JudgeResult {
  double pos = 0.0, ratio[3] = {0.0, 0.0, 0.0};
  uint32_t body = 0;
};
mfree_status judge(mfree_surrogate* s, mfree_ctx* c, JudgeResult* r) {
  const int b = c->cur;
  constexpr int kS = kJudgeSums * kJudgeGroups;
  hipLaunchKernelGGL(k_judge_partial, dim3(c->nchunks), dim3(256), 0, c->stream, c->chunks, c->N, c->y[b],
                     c->gid[b], s->jstart, s->cand, s->base, s->devp);
  std::vector<double> h(kS * (size_t)c->nchunks);
  SCHK(s, hipMemcpyAsync(h.data(), s->devp, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  SCHK(s, hipStreamSynchronize(c->stream));
  *r = JudgeResult();
  double worst = -1.0;
  // chunks were built body by body, kChunk particles each
  uint32_t ch = 0;
  for (uint32_t bd = 0; bd < c->S; bd++) {
    double t[kS] = {};
    for (uint32_t o = c->off[bd]; o < c->off[bd + 1]; o += kChunk, ch++)
      for (int k = 0; k < kS; k++) t[k] += h[ch * kS + k];
    const double dpos = t[2] + 1e-12 * t[3];
    const double pos = dpos > 0.0 ? sqrt(t[0] / dpos) : (t[0] > 0.0 ? INFINITY : 0.0);
    double bad = pos;
    if (!(pos <= r->pos)) r->pos = pos;
    const SceneConst& k = c->h_sc[bd];
    const double scale[3] = {sqrt(k.c0sq), k.rho0, k.K};
    const double np = (double)(c->off[bd + 1] - c->off[bd]);
    for (int q = 1; q < kJudgeGroups; q++) {
      const double nl = t[kJudgeSums * q], nb = t[kJudgeSums * q + 1];
      const double fl = kNetFloor * scale[q - 1];
      const double den = nb + fl * fl * np;
      const double ratio = den > 0.0 ? sqrt(nl / den) : (nl > 0.0 ? INFINITY : 0.0);
      if (!(ratio <= r->ratio[q - 1])) r->ratio[q - 1] = ratio;
      if (!(ratio <= bad)) bad = ratio;
    }
    if (!(bad <= worst)) {
      worst = bad;
      r->body = bd;
    }
  }
  return MFREE_OK;
}

void copy_async(void* dst, const void* src, size_t bytes, hipStream_t st) {
  (void)hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, st);
}

// Snapshot of the full per-particle storage of the current buffer (rollback).
void snapshot(mfree_surrogate* s, mfree_ctx* c, std::vector<double>& times) {
  const int b = c->cur;
  const size_t N = c->N;
  for (int k = 0; k < 8; k++) copy_async(s->snap + k * N, c->y[b].f[k], N * 8, c->stream);
  copy_async(s->snap + 8 * N, c->m[b], N * 8, c->stream);
  copy_async(s->snap + 9 * N, c->h[b], N * 8, c->stream);
  copy_async(s->snapu, c->scene[b], N * 4, c->stream);
  copy_async(s->snapu + N, c->gid[b], N * 4, c->stream);
  copy_async(s->snap_sc, c->sc, c->S * sizeof(SceneConst), c->stream);
  times = c->time;
}

void restore(mfree_surrogate* s, mfree_ctx* c, const std::vector<double>& times) {
  const int b = c->cur;
  const size_t N = c->N;
  for (int k = 0; k < 8; k++) copy_async(c->y[b].f[k], s->snap + k * N, N * 8, c->stream);
  copy_async(c->m[b], s->snap + 8 * N, N * 8, c->stream);
  copy_async(c->h[b], s->snap + 9 * N, N * 8, c->stream);
  copy_async(c->scene[b], s->snapu, N * 4, c->stream);
  copy_async(c->gid[b], s->snapu + N, N * 4, c->stream);
  copy_async(c->sc, s->snap_sc, c->S * sizeof(SceneConst), c->stream);
  c->time = times;
}

double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace
}  // namespace mfree

// ---------------------------------------------------------------------------

extern "C" {

void mfree_surrogate_default_options(mfree_surrogate_options* o) {
  if (!o) return;
  memset(o, 0, sizeof(*o));
  o->device = 0;
  o->seed = 1;
  o->max_batch_nodes = 1u << 20;
  o->max_compression = 1.0f;
  o->base_order = 4;
}

mfree_status mfree_surrogate_create(const mfree_surrogate_options* o, mfree_surrogate** out) {
  if (!out) return MFREE_ERR_ARG;
  *out = nullptr;
  mfree_surrogate* s = new mfree_surrogate();
  if (o)
    s->opt = *o;
  else
    mfree_surrogate_default_options(&s->opt);
  if (s->opt.base_order != 3 && s->opt.base_order != 4) {
    delete s;
    return MFREE_ERR_ARG;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= s->opt.device) {
    delete s;
    return MFREE_ERR_HIP;
  }
  hipDeviceProp_t prop;
  if (hipSetDevice(s->opt.device) != hipSuccess || hipGetDeviceProperties(&prop, s->opt.device) != hipSuccess ||
      strncmp(prop.gcnArchName, "gfx950", 6) != 0 || hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking) != hipSuccess) {
    delete s;
    return MFREE_ERR_HIP;
  }
  s->rng.seed(s->opt.seed);
  std::string msg;
  if (mp_net_init(s->net, s->opt.seed, msg) != MFREE_OK || hipMalloc(&s->dnorm, kNorm * sizeof(float)) != hipSuccess) {
    mp_net_free(s->net);
    hipStreamDestroy(s->st);
    delete s;
    return MFREE_ERR_HIP;
  }
  for (int k = 0; k < kNorm; k++) s->hnorm[k] = 0.f;
  for (int k = 0; k < kNF; k++) s->hnorm[kNF + k] = 1.f;
  for (int k = 0; k < kOut; k++) s->hnorm[2 * kNF + kOut + k] = 1.f;
  (void)hipMemcpy(s->dnorm, s->hnorm, sizeof(s->hnorm), hipMemcpyHostToDevice);
  *out = s;
  return MFREE_OK;
}

void mfree_surrogate_destroy(mfree_surrogate* s) {
  if (!s) return;
  (void)hipSetDevice(s->opt.device);
  (void)hipDeviceSynchronize();
  drop_graphs(s);
  mp_net_free(s->net);
  mp_acts_free(s->acts);
  mp_graph_free(s->graph);
  free_list(s->wallocs);
  free_list(s->tallocs);
  free_list(s->callocs);
  (void)hipFree(s->dnorm);
  (void)hipFree(s->dsF);
  (void)hipFree(s->dsT);
  (void)hipFree(s->dsX);
  (void)hipFree(s->dsY);
  (void)hipFree(s->dsH);
  (void)hipFree(s->loss);
  if (s->st) (void)hipStreamDestroy(s->st);
  delete s;
}

const char* mfree_surrogate_last_error(const mfree_surrogate* s) { return s ? s->err.c_str() : "null surrogate"; }

uint64_t mfree_mpnn_param_count(void) {
  return (uint64_t)kH * kNF + kH + kLayers * ((uint64_t)kC * kH + (uint64_t)kH * kG + kH) + (uint64_t)kOut * kH + kOut;
}

uint32_t mfree_mpnn_layout(int32_t* shapes) {
  const int32_t t[] = {kH, kNF, 1, kH, kC, kH, kH, kG, 1, kH, kC, kH, kH, kG, 1, kH,
                       kC, kH, kH, kG, 1, kH, kOut, kH, 1, kOut};
  const uint32_t nt = (uint32_t)(sizeof(t) / sizeof(t[0]) / 2);
  if (shapes) memcpy(shapes, t, sizeof(t));
  return nt;
}

mfree_status mfree_surrogate_get_params(mfree_surrogate* s, float* w, float* norm) {
  if (!s) return MFREE_ERR_ARG;
  (void)hipSetDevice(s->opt.device);
  if (w) SCHK(s, hipMemcpy(w, s->net.w, s->net.nparam * sizeof(float), hipMemcpyDeviceToHost));
  if (norm) memcpy(norm, s->hnorm, sizeof(s->hnorm));
  return MFREE_OK;
}

mfree_status mfree_surrogate_set_params(mfree_surrogate* s, const float* w, const float* norm) {
  if (!s) return MFREE_ERR_ARG;
  (void)hipSetDevice(s->opt.device);
  SCHK(s, hipDeviceSynchronize());
  if (w) {
    SCHK(s, hipMemcpy(s->net.w, w, s->net.nparam * sizeof(float), hipMemcpyHostToDevice));
    mp_net_sync16(s->net, s->st);
    SCHK(s, hipStreamSynchronize(s->st));
  }
  if (norm) {
    for (int k = 0; k < kNorm; k++)
      if (!isfinite(norm[k])) return sfail(s, MFREE_ERR_ARG, "non-finite normalisation");
    for (int k = 0; k < kNF; k++)
      if (!(norm[kNF + k] > 0.f)) return sfail(s, MFREE_ERR_ARG, "feature std must be positive");
    memcpy(s->hnorm, norm, sizeof(s->hnorm));
    SCHK(s, hipMemcpy(s->dnorm, s->hnorm, sizeof(s->hnorm), hipMemcpyHostToDevice));
    s->have_norm = true;
  }
  return MFREE_OK;
}

uint64_t mfree_surrogate_snapshots(const mfree_surrogate* s) {
  if (!s) return 0;
  uint64_t n = 0;
  for (const auto& c : s->caps)
    for (uint8_t v : c.valid) n += v;
  return n;
}

uint64_t mfree_surrogate_dataset_nodes(const mfree_surrogate* s) { return s ? s->rows : 0; }

mfree_status mfree_surrogate_collect(mfree_surrogate* s, mfree_ctx* c, uint32_t ncapture, uint32_t gap) {
  if (!s || !c) return MFREE_ERR_ARG;
  SCHK(s, hipSetDevice(s->opt.device));
  mfree_status st = bind_ctx(s, c);
  if (st != MFREE_OK) return st;
  const uint64_t N = c->N;
  st = ensure_rows(s, s->rows + (uint64_t)ncapture * N);
  if (st != MFREE_OK) return st;
  for (uint32_t k = 0; k < ncapture; k++) {
    // the velocity one learned step before the capture: recorded here at the
    // first capture or after a gap, else before the previous capture's
    // target steps
    if (k == 0 || gap) {
      if (gap && k > 0) {
        st = run_steps(c, gap);
        if (st != MFREE_OK) return sfail(s, st, std::string("collect: ") + c->err);
      }
      record_vprev(s, c);
      st = run_steps(c, kStepsPerLearned);
      if (st != MFREE_OK) return sfail(s, st, std::string("collect: ") + c->err);
    }
    // inputs and base update at the capture state
    const uint64_t row0 = s->rows;
    launch_fill_u32(c->stream, s->comp, 0u, c->S);
    hipLaunchKernelGGL(k_body_compression, dim3(nblk(N, 256)), dim3(256), 0, c->stream, (uint32_t)N,
                       c->y[c->cur].f[4], c->scene[c->cur], c->sc, s->comp);
    // the base step writes the other storage buffer and leaves the capture
    // state where it is
    const int a0 = c->cur;
    enqueue_base(s, c, true, false);
    StepArgs a = step_args(s, c);
    a.F = s->dsF + row0 * kNF;
    a.base = s->base;
    a.inv = s->inv;
    a.ox = s->dsX + row0;
    a.oy = s->dsY + row0;
    a.oh = s->dsH + row0;
    hipLaunchKernelGGL(k_learned, dim3(nblk(N, 256)), dim3(256), 0, c->stream, a);
    c->cur = a0;
    st = sync_ctx(s, c);
    if (st != MFREE_OK) return st;
    record_vprev(s, c);
    // truth: 4 RK4 steps
    st = run_steps(c, kStepsPerLearned);
    if (st != MFREE_OK) return sfail(s, st, std::string("collect: ") + c->err);
    const int b = c->cur;
    hipLaunchKernelGGL(k_targets, dim3(nblk(N, 256)), dim3(256), 0, c->stream, (uint32_t)N, c->y[b], c->gid[b],
                       c->scene[b], c->sc, c->h[b], s->base, s->inv, s->dsT + row0 * kOut);
    hipLaunchKernelGGL(k_body_compression, dim3(nblk(N, 256)), dim3(256), 0, c->stream, (uint32_t)N, c->y[b].f[4],
                       c->scene[b], c->sc, s->comp);
    std::vector<float> comp(c->S);
    SCHK(s, hipMemcpyAsync(comp.data(), s->comp, c->S * 4, hipMemcpyDeviceToHost, c->stream));
    SCHK(s, hipStreamSynchronize(c->stream));
    mfree_surrogate::Capture cp;
    cp.row0 = row0;
    cp.off = c->off;
    cp.valid.resize(c->S);
    // a body compressed beyond max_compression (|rho/rho0 - 1|) at either
    // end of the step is no elastic-solid sample: not used for training
    for (uint32_t k = 0; k < c->S; k++) cp.valid[k] = comp[k] <= s->opt.max_compression ? 1 : 0;
    s->caps.push_back(cp);
    s->rows += N;
  }
  s->wins_ready = false;
  return compute_norm(s);
}

mfree_status mfree_surrogate_train(mfree_surrogate* s, uint32_t iters, float lr, float* loss) {
  if (!s || !(lr > 0.f)) return MFREE_ERR_ARG;
  SCHK(s, hipSetDevice(s->opt.device));
  mfree_status st = build_windows(s);
  if (st != MFREE_OK) return st;
  if (iters > s->loss_cap) {
    (void)hipFree(s->loss);
    s->loss = nullptr;
    SCHK(s, hipMalloc(&s->loss, iters * sizeof(float)));
    s->loss_cap = iters;
  }
  if (iters) SCHK(s, hipMemsetAsync(s->loss, 0, iters * sizeof(float), s->st));
  std::uniform_int_distribution<size_t> pick(0, s->wins.size() - 1);
  const float lr_min = 0.05f * lr;
  for (uint32_t it = 0; it < iters; it++) {
    // cosine decay to 5% of the initial rate
    const float t = iters > 1 ? (float)it / (float)(iters - 1) : 0.f;
    const float lr_t = lr_min + 0.5f * (lr - lr_min) * (1.f + cosf(3.14159265f * t));
    enqueue_train_step(s, s->wins[pick(s->rng)], lr_t, s->loss + it);
  }
  SCHK(s, hipStreamSynchronize(s->st));
  SCHK(s, hipGetLastError());
  uint32_t e = 0;
  SCHK(s, hipMemcpy(&e, s->terr, sizeof(uint32_t), hipMemcpyDeviceToHost));
  if (e) {
    uint32_t z = 0;
    SCHK(s, hipMemcpy(s->terr, &z, sizeof(uint32_t), hipMemcpyHostToDevice));
    if (e & EB_NBH) return sfail(s, MFREE_ERR_NBH_OVERFLOW, "training graph: a node has more than 64 edges");
    return sfail(s, MFREE_ERR_CELLS, "training binning failed");
  }
  if (loss && iters) SCHK(s, hipMemcpy(loss, s->loss, iters * sizeof(float), hipMemcpyDeviceToHost));
  return MFREE_OK;
}

mfree_status mfree_surrogate_step(mfree_surrogate* s, mfree_ctx* c, int32_t use_net) {
  if (!s || !c) return MFREE_ERR_ARG;
  SCHK(s, hipSetDevice(s->opt.device));
  mfree_status st = bind_ctx(s, c);
  if (st != MFREE_OK) return st;
  if (use_net && !s->have_norm) return sfail(s, MFREE_ERR_ARG, "network has no normalisation (collect or set_params)");
  if (!s->vprev_valid) record_vprev(s, c);
  st = enqueue_learned(s, c, use_net != 0);
  if (st != MFREE_OK) return st;
  return sync_ctx(s, c);
}

mfree_status mfree_surrogate_record_velocity(mfree_surrogate* s, mfree_ctx* c) {
  if (!s || !c) return MFREE_ERR_ARG;
  SCHK(s, hipSetDevice(s->opt.device));
  mfree_status st = bind_ctx(s, c);
  if (st != MFREE_OK) return st;
  record_vprev(s, c);
  SCHK(s, hipStreamSynchronize(c->stream));
  return MFREE_OK;
}

mfree_status mfree_surrogate_rollout(mfree_surrogate* s, mfree_ctx* c, uint32_t nsteps, uint32_t group,
                                     double threshold, mfree_judge_stats* stats) {
  if (!s || !c || group == 0) return MFREE_ERR_ARG;
  SCHK(s, hipSetDevice(s->opt.device));
  mfree_status st = bind_ctx(s, c);
  if (st != MFREE_OK) return st;
  if (!s->have_norm) return sfail(s, MFREE_ERR_ARG, "network has no normalisation (collect or set_params)");
  mfree_judge_stats js;
  memset(&js, 0, sizeof(js));
  uint32_t done = 0;
  auto plain = [&](uint32_t k) -> mfree_status {
    mfree_status r = run_steps(c, k);
    js.plain_steps += k;
    return r == MFREE_OK ? MFREE_OK : sfail(s, r, std::string("plain steps: ") + c->err);
  };
  // warm start: the features need the velocity one learned step back
  if (nsteps >= 3 * kStepsPerLearned) {
    record_vprev(s, c);
    st = plain(kStepsPerLearned);
    if (st != MFREE_OK) return st;
    done += kStepsPerLearned;
  }
  std::vector<double> times;
  const size_t N = c->N;
  // a numerical failure of the learned steps (not of the plain ones) makes
  // the group fall back like a rejected one
  auto numeric = [](mfree_status r) {
    return r == MFREE_ERR_NONFINITE || r == MFREE_ERR_NBH_OVERFLOW || r == MFREE_ERR_NO_NEIGHBOR ||
           r == MFREE_ERR_CELLS || r == MFREE_ERR_SKIN;
  };
  // groups of `group` learned steps; the last one may be shorter (>= 2)
  while (nsteps - done >= 2 * kStepsPerLearned) {
    const uint32_t glen = std::min(group, (nsteps - done) / kStepsPerLearned);
    const uint32_t gspan = glen * kStepsPerLearned;
    auto t0 = std::chrono::steady_clock::now();
    snapshot(s, c, times);
    for (uint32_t k = 0; k + 1 < glen; k++) {
      st = enqueue_learned(s, c, true);
      if (st != MFREE_OK) return st;
    }
    st = sync_ctx(s, c);
    bool failed = numeric(st);
    if (st != MFREE_OK && !failed) return st;
    js.ms_learned += ms_since(t0);
    auto t1 = std::chrono::steady_clock::now();
    JudgeResult jr;
    if (!failed) {
      // judge: the last learned step (its base step kept) and 4 plain steps
      // from the same state (the learned step writes the other storage
      // buffer: the judged state stays in this one)
      const int a0 = c->cur;
      copy_async(s->jsc, c->sc, c->S * sizeof(SceneConst), c->stream);
      const std::vector<double> jtimes = c->time;
      hipLaunchKernelGGL(k_state_by_gid, dim3(nblk(N, 256)), dim3(256), 0, c->stream, (uint32_t)N, c->y[a0],
                         c->gid[a0], s->jstart);
      st = enqueue_learned(s, c, true, s->base);  // also records the velocities of this state
      if (st != MFREE_OK) return st;
      const int b = c->cur;
      hipLaunchKernelGGL(k_state_by_gid, dim3(nblk(N, 256)), dim3(256), 0, c->stream, (uint32_t)N, c->y[b],
                         c->gid[b], s->cand);
      c->cur = a0;
      copy_async(c->sc, s->jsc, c->S * sizeof(SceneConst), c->stream);
      c->time = jtimes;
      st = sync_ctx(s, c);
      failed = numeric(st);
      if (st != MFREE_OK && !failed) return st;
    }
    if (!failed) {
      st = plain(kStepsPerLearned);
      if (st != MFREE_OK) return st;
      st = judge(s, c, &jr);
      if (st != MFREE_OK) return st;
      js.ms_judge += ms_since(t1);
      js.last_deviation = jr.pos;
      if (!(jr.pos <= js.max_deviation)) js.max_deviation = jr.pos;
      js.max_ratio_vel = std::max(js.max_ratio_vel, jr.ratio[0]);
      js.max_ratio_rho = std::max(js.max_ratio_rho, jr.ratio[1]);
      js.max_ratio_stress = std::max(js.max_ratio_stress, jr.ratio[2]);
    }
    js.groups++;
    // (an infinite threshold accepts every group that completes)
    const bool ratios_ok = std::isinf(threshold) && threshold > 0
                               ? true
                               : jr.ratio[0] <= kNetRatioLimit && jr.ratio[1] <= kNetRatioLimit &&
                                     jr.ratio[2] <= kNetRatioLimit;
    const bool accept = !failed && jr.pos <= threshold && ratios_ok;
    if (accept) {
      // kept: the group's learned steps but the judged one, whose 4 plain
      // steps are counted in plain_steps
      js.learned_steps += glen - 1;
    } else {
      // roll back the group and redo it with plain RK4 steps (the judge's
      // plain steps are discarded with it)
      auto t2 = std::chrono::steady_clock::now();
      js.fallbacks++;
      if (failed) {
        js.numeric_fallbacks++;
      } else {
        js.worst_body = jr.body;
        js.plain_steps -= kStepsPerLearned;
      }
      restore(s, c, times);
      st = plain(gspan - kStepsPerLearned);
      if (st != MFREE_OK) return st;
      record_vprev(s, c);
      st = plain(kStepsPerLearned);
      if (st != MFREE_OK) return st;
      js.ms_fallback += ms_since(t2);
    }
    done += gspan;
  }
  if (done < nsteps) {
    if (nsteps - done >= kStepsPerLearned) {
      st = plain(nsteps - done - kStepsPerLearned);
      if (st != MFREE_OK) return st;
      record_vprev(s, c);
      st = plain(kStepsPerLearned);
    } else {
      st = plain(nsteps - done);
    }
    if (st != MFREE_OK) return st;
  }
  if (stats) *stats = js;
  return MFREE_OK;
}

mfree_status mfree_surrogate_graph(mfree_surrogate* s, mfree_ctx* c, uint32_t* row_ptr, uint32_t* col, float* edge3,
                                   uint64_t cap, uint64_t* nedges) {
  if (!s || !c || !row_ptr) return MFREE_ERR_ARG;
  SCHK(s, hipSetDevice(s->opt.device));
  mfree_status st = bind_ctx(s, c);
  if (st != MFREE_OK) return st;
  // as in a learned step (the base step builds the graph; the state stays in
  // its storage buffer)
  const int a0 = c->cur;
  enqueue_base(s, c, true, false);
  const int b = c->cur;
  c->cur = a0;
  st = sync_ctx(s, c);
  if (st != MFREE_OK) return st;
  const size_t N = c->N;
  const size_t stride = s->graph.cap;  // >= N
  std::vector<uint32_t> deg(N), gid(N), colv(stride * kMaxDeg);
  SCHK(s, hipMemcpy(deg.data(), s->graph.deg, N * 4, hipMemcpyDeviceToHost));
  SCHK(s, hipMemcpy(gid.data(), c->gid[b], N * 4, hipMemcpyDeviceToHost));  // node -> global id
  SCHK(s, hipMemcpy(colv.data(), s->graph.col, stride * kMaxDeg * 4, hipMemcpyDeviceToHost));
  std::vector<uint32_t> rp(N + 1, 0);
  for (size_t p = 0; p < N; p++) rp[p + 1] = rp[p] + deg[p];
  const uint64_t tot = rp[N];
  if (nedges) *nedges = tot;
  if (tot > cap) return sfail(s, MFREE_ERR_CAPACITY, "edge capacity too small");
  std::vector<float> ew(3 * std::max<uint64_t>(tot, 1));
  if (edge3 && tot) {
    float* de = nullptr;
    uint32_t* drp = nullptr;
    SCHK(s, hipMalloc(&de, tot * 3 * sizeof(float)));
    if (hipMalloc(&drp, (N + 1) * 4) != hipSuccess) {
      (void)hipFree(de);
      return sfail(s, MFREE_ERR_HIP, "hipMalloc failed");
    }
    hipError_t e = hipMemcpy(drp, rp.data(), (N + 1) * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_edge_export, dim3(nblk(N, 256)), dim3(256), 0, c->stream, (uint32_t)N, (uint32_t)stride,
                         s->graph.deg, s->graph.col, drp, s->graph.npos, de);
      e = hipStreamSynchronize(c->stream);
    }
    if (e == hipSuccess) e = hipMemcpy(ew.data(), de, tot * 3 * sizeof(float), hipMemcpyDeviceToHost);
    (void)hipFree(de);
    (void)hipFree(drp);
    if (e != hipSuccess) return sfail(s, MFREE_ERR_HIP, hipGetErrorString(e));
  }
  // rows by global id, neighbours as global ids in node order
  std::vector<uint32_t> row_of(N);
  for (size_t p = 0; p < N; p++) row_of[gid[p]] = (uint32_t)p;
  row_ptr[0] = 0;
  for (size_t g = 0; g < N; g++) row_ptr[g + 1] = row_ptr[g] + deg[row_of[g]];
  for (size_t g = 0; g < N; g++) {
    const size_t p = row_of[g];
    for (uint32_t k = 0; k < deg[p]; k++) {
      const size_t o = row_ptr[g] + k, e = rp[p] + k;
      if (col) col[o] = gid[colv[(size_t)k * stride + p]];
      if (edge3) {
        edge3[3 * o] = ew[3 * e];
        edge3[3 * o + 1] = ew[3 * e + 1];
        edge3[3 * o + 2] = ew[3 * e + 2];
      }
    }
  }
  return MFREE_OK;
}

mfree_status mfree_mpnn_forward_backward(int32_t device, uint32_t n, const uint32_t* row_ptr, const uint32_t* col,
                                         const float* pos3, const float* x, const float* t, const float* w,
                                         float* y, float* loss, float* grad) {
  if (!row_ptr || !col || !pos3 || !x || !w || n == 0 || (grad && !t)) return MFREE_ERR_ARG;
  const bool train = grad != nullptr;  // without grad: the inference forward only
  if (row_ptr[0] != 0) return MFREE_ERR_ARG;
  for (uint32_t i = 0; i < n; i++) {
    if (row_ptr[i + 1] < row_ptr[i] || row_ptr[i + 1] - row_ptr[i] > kMaxDeg) return MFREE_ERR_ARG;
    for (uint32_t e = row_ptr[i]; e < row_ptr[i + 1]; e++)
      if (col[e] >= n) return MFREE_ERR_ARG;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || hipSetDevice(device) != hipSuccess)
    return MFREE_ERR_HIP;
  MpNet net;
  MpActs acts;
  MpGraph g;
  std::string msg;
  mfree_status st = mp_net_init(net, 0, msg);
  if (st == MFREE_OK) st = mp_acts_reserve(acts, n, train, msg);
  if (st == MFREE_OK) st = mp_graph_reserve(g, n, msg);
  float *dT = nullptr, *dL = nullptr;
  hipStream_t strm = nullptr;
  if (st == MFREE_OK && (hipMalloc(&dT, (size_t)n * kOut * 4) != hipSuccess || hipMalloc(&dL, 4) != hipSuccess ||
                         hipStreamCreateWithFlags(&strm, hipStreamNonBlocking) != hipSuccess))
    st = MFREE_ERR_HIP;
  if (st == MFREE_OK) {
    std::vector<float4> np(n);
    std::vector<uint32_t> deg(n), cs((size_t)n * kMaxDeg, 0);
    std::vector<bf16_t> x16((size_t)n * kK0, 0);
    for (uint32_t i = 0; i < n; i++) {
      np[i] = make_float4(pos3[3 * (size_t)i], pos3[3 * (size_t)i + 1], pos3[3 * (size_t)i + 2], 0.f);
      deg[i] = row_ptr[i + 1] - row_ptr[i];
      for (uint32_t k = 0; k < deg[i]; k++) cs[(size_t)k * n + i] = col[row_ptr[i] + k];
      for (int k = 0; k < kNF; k++) x16[(size_t)i * kK0 + k] = f2bf(x[(size_t)i * kNF + k]);
    }
    const float zero = 0.f;
    bool ok = hipMemcpy(g.deg, deg.data(), (size_t)n * 4, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(g.col, cs.data(), cs.size() * 4, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(g.npos, np.data(), (size_t)n * 16, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(acts.X0, x16.data(), x16.size() * 2, hipMemcpyHostToDevice) == hipSuccess &&
              (!train || hipMemcpy(dT, t, (size_t)n * kOut * 4, hipMemcpyHostToDevice) == hipSuccess) &&
              hipMemcpy(dL, &zero, 4, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(net.w, w, net.nparam * 4, hipMemcpyHostToDevice) == hipSuccess;
    if (ok) {
      mp_net_sync16(net, strm);
      mp_forward(net, acts, g, n, train, strm);
      if (train) {
        mp_zero_grad(net, strm);
        mp_backward(net, acts, g, n, dT, dL, strm);
      }
      ok = hipStreamSynchronize(strm) == hipSuccess;
    }
    if (ok && y) {
      std::vector<float> yy((size_t)n * kOutP);
      ok = hipMemcpy(yy.data(), acts.Y, yy.size() * 4, hipMemcpyDeviceToHost) == hipSuccess;
      for (uint32_t i = 0; ok && i < n; i++)
        for (int k = 0; k < kOut; k++) y[(size_t)i * kOut + k] = yy[(size_t)i * kOutP + k];
    }
    if (ok && loss && train) ok = hipMemcpy(loss, dL, 4, hipMemcpyDeviceToHost) == hipSuccess;
    if (ok && grad) ok = hipMemcpy(grad, net.g, net.nparam * 4, hipMemcpyDeviceToHost) == hipSuccess;
    if (!ok) st = MFREE_ERR_HIP;
  }
  if (strm) (void)hipStreamDestroy(strm);
  (void)hipFree(dT);
  (void)hipFree(dL);
  mp_graph_free(g);
  mp_acts_free(acts);
  mp_net_free(net);
  return st;
}

}  // extern "C"
