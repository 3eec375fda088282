// FP64 energy of a body (the path's energy-conservation check) and its
// allocation-free variant.
//
//   kinetic      sum m |v|^2 / 2
//   volumetric   sum m c0^2 (ln(rho/rho0) + rho0/rho - 1): the work stored by
//                the linear EOS p = c0^2 (rho - rho0) (src/material.h:37-64),
//                since d/drho of it is p / rho^2
//   deviatoric   sum (m/rho) S:S / (4G), S:S = Sxx^2 + 2 Sxy^2 + Syy^2
//
// One workgroup sums each chunk of kEnergyChunk consecutive particles into a
// partial, and one thread adds the partials in chunk order: the result
// depends only on the particles' order, not on scheduling or grid size, so
// both entry points give bit-identical sums for the same arrays.
#include "../../include/mfree_iwf_amd.h"

#include <math.h>

#include <mutex>

#include "ctx.h"

namespace mfree {

constexpr uint32_t kEnergyChunk = 4096;   // particles per partial (256 threads x 16)
// Allocation-free entry point: partials live in the file-scope array below.
constexpr uint32_t kScratchChunks = 4096;  // -> at most 16,777,216 particles per call
constexpr uint64_t kScratchMaxParticles = (uint64_t)kScratchChunks * kEnergyChunk;

struct EnergyIn {
  const double *m, *rho, *vx, *vy, *Sxx, *Sxy, *Syy;
  uint32_t n;
  double rho0, c0sq, G2;
};

__global__ __launch_bounds__(256) void k_energy_chunks(EnergyIn in, double* __restrict__ part) {
  const uint32_t b = blockIdx.x * kEnergyChunk, e = min(b + kEnergyChunk, in.n);
  double e0 = 0.0, e1 = 0.0, e2 = 0.0;
  for (uint32_t p = b + threadIdx.x; p < e; p += 256u) {
    const double mi = in.m[p], vx = in.vx[p], vy = in.vy[p], rho = in.rho[p];
    const double sxx = in.Sxx[p], sxy = in.Sxy[p], syy = in.Syy[p];
    e0 += 0.5 * mi * (vx * vx + vy * vy);
    const double d = (rho - in.rho0) / in.rho0;
    e1 += mi * in.c0sq * (log1p(d) - d / (1.0 + d));
    e2 += mi / rho * (sxx * sxx + 2.0 * sxy * sxy + syy * syy) / (2.0 * in.G2);
  }
  __shared__ double red[3][256];
  red[0][threadIdx.x] = e0;
  red[1][threadIdx.x] = e1;
  red[2][threadIdx.x] = e2;
  __syncthreads();
  for (uint32_t w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int k = 0; k < 3; k++) red[k][threadIdx.x] += red[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 3) part[blockIdx.x * 3 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void k_energy_final(const double* __restrict__ part, uint32_t nchunks, double* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double e[3] = {0.0, 0.0, 0.0};
  for (uint32_t b = 0; b < nchunks; b++)
    for (int k = 0; k < 3; k++) e[k] += part[b * 3 + k];
  for (int k = 0; k < 3; k++) out[k] = e[k];
}

static void launch_energy(hipStream_t st, const EnergyIn& in, double* part, double* out) {
  const uint32_t nchunks = (in.n + kEnergyChunk - 1) / kEnergyChunk;
  hipLaunchKernelGGL(k_energy_chunks, dim3(nchunks), dim3(256), 0, st, in, part);
  hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(1), 0, st, part, nchunks, out);
}

// Scratch of the allocation-free entry point: kScratchChunks partials (3
// doubles each, 96 KiB in all).  Each GPU the library runs on holds its own
// copy until the process ends; every call on one GPU shares it, so calls
// that overlap on one GPU race.  Host code never passes this name to a
// kernel: it uses the address hipGetSymbolAddress returns for the device.
static __device__ __align__(256) double scratch[kScratchChunks * 3];

}  // namespace mfree

using namespace mfree;

extern "C" {

mfree_status mfree_energy(mfree_ctx* c, uint32_t body, double* energy) {
  if (!c || body >= c->prm.size() || !energy) return MFREE_ERR_ARG;
  mfree_status st = commit(c);
  if (st) return st;
  HIPCHK(c, hipSetDevice(c->opt.device));
  const uint32_t o = c->off[body], n = c->off[body + 1] - o;
  if (!c->energy_buf) {
    // partials of the largest body, then the 3 results
    uint32_t nmax = 0;
    for (size_t s = 0; s + 1 < c->off.size(); s++) nmax = std::max(nmax, c->off[s + 1] - c->off[s]);
    const size_t nchunks = (nmax + kEnergyChunk - 1) / kEnergyChunk;
    c->energy_out = nchunks * 3;
    HIPCHK(c, hipMalloc(&c->energy_buf, (c->energy_out + 3) * sizeof(double)));
    c->allocs.push_back(c->energy_buf);
  }
  const int b = c->cur;
  const SceneConst& k = c->h_sc[body];
  EnergyIn in{c->m[b] + o,         c->y[b].f[4] + o, c->y[b].f[2] + o, c->y[b].f[3] + o, c->y[b].f[5] + o,
              c->y[b].f[6] + o,    c->y[b].f[7] + o, n,                k.rho0,           k.c0sq,
              k.G2};
  launch_energy(c->stream, in, c->energy_buf, c->energy_buf + c->energy_out);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(energy, c->energy_buf + c->energy_out, 3 * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MFREE_OK;
}

mfree_status mfree_energy_noalloc(int32_t device, void* stream, const mfree_params* p, const mfree_particles* dev,
                                  uint32_t n, double* energy_dev) {
  if (!p || !dev || !energy_dev || !dev->m || !dev->rho || !dev->vx || !dev->vy || !dev->Sxx || !dev->Sxy ||
      !dev->Syy || n == 0 || !(p->rho0 > 0))
    return MFREE_ERR_ARG;
  if ((uint64_t)n > kScratchMaxParticles) return MFREE_ERR_SCRATCH;
  constexpr int kMaxDevices = 64;
  static double* dev_scratch[kMaxDevices] = {};
  static std::mutex mu;
  if (device < 0 || device >= kMaxDevices) return MFREE_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return MFREE_ERR_HIP;
  double* part = nullptr;
  {
    std::lock_guard<std::mutex> lk(mu);
    if (!dev_scratch[device]) {
      void* q = nullptr;
      if (hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)) != hipSuccess || !q) return MFREE_ERR_HIP;
      dev_scratch[device] = (double*)q;
    }
    part = dev_scratch[device];
  }
  SceneConst k;
  scene_consts(*p, 0, n, k);
  EnergyIn in{dev->m, dev->rho, dev->vx, dev->vy, dev->Sxx, dev->Sxy, dev->Syy, n, k.rho0, k.c0sq, k.G2};
  launch_energy((hipStream_t)stream, in, part, energy_dev);
  return hipGetLastError() == hipSuccess ? MFREE_OK : MFREE_ERR_HIP;
}

}  // extern "C"
