// Internal types shared by the host orchestration (capi.hip) and the
// gfx950 kernels (grid.hip, ul.hip, tl.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mfree_iwf_amd.h"

namespace mfree {

constexpr int kBlock = 256;          // threads per workgroup (4 wave64)
constexpr uint32_t kChunk = 2048;    // particles per bbox-reduction chunk
constexpr uint32_t kScanTile = 1024; // elements per scan workgroup

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  Kernels whose neighbouring workgroups share data remap the block
// index so that XCD x processes the contiguous range of tiles
// [x*per, (x+1)*per); the grid is padded to a multiple of 8 and blocks
// mapped past `ntiles` exit.
constexpr uint32_t kXcds = 8;
__host__ __device__ inline uint32_t xcd_grid(uint32_t ntiles) { return (ntiles + kXcds - 1) / kXcds * kXcds; }
__device__ inline uint32_t xcd_tile() {
  const uint32_t per = gridDim.x / kXcds;
  return (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
}
// Default XCD chunk of the per-particle pair kernels: 64 tiles of 256
// particles, about 16 cell columns of a 1M-particle ring scene (neighbour
// reads stay mostly within one XCD's chunk) while a dense stretch of tiles
// (a contact zone) is shared by several XCDs.
constexpr uint32_t kXcdChunkDefault = 64;
// Neighbour-list path: slots per particle, the skin's floor in units of h_max
// and its multiple of dt max|v| (twice the largest pair approach of a step at
// twice the fastest particle's speed).  A list reaches 2 h + skin: about 30
// entries at rest for slow bodies, up to about 85 at the largest skin (2 h_max,
// Mach > 2 impacts), so 512 slots (2 KiB per particle) hold a 6-fold
// compression before the list overflows (EB_NBH).
constexpr uint32_t kListCap = 512;
// mixed pair passes: a candidate word holds the sorted position in its low
// 28 bits and the cell offsets (dx + 1, dy + 1) in bits 28-29 and 30-31
constexpr uint32_t kListMask = 0x0FFFFFFFu;
constexpr double kSkinMin = 0.25;
constexpr double kSkinVfac = 4.0;
// a skin above kSkinMax h_max (a particle moving h/2 in one step) is no
// stable step: the body's binning fails (EB_SKIN) instead of degenerating
// into one cell
constexpr double kSkinMax = 2.0;
// The same with the tiles dealt to the XCDs in chunks of `chunk`
// consecutive tiles (XCD x takes chunks x, x + 8, ...; the remainder is
// split evenly), so that a costly stretch of tiles is shared by several
// XCDs; chunk = 0: one contiguous range per XCD (xcd_tile).  A bijection
// of [0, gridDim.x) when gridDim.x is a multiple of kXcds.
__device__ inline uint32_t xcd_tile_chunked(uint32_t chunk) {
  const uint32_t per = gridDim.x / kXcds;
  const uint32_t x = blockIdx.x % kXcds, j = blockIdx.x / kXcds;
  if (chunk == 0 || chunk >= per) return x * per + j;
  const uint32_t full = per / chunk, c = j / chunk, o = j - c * chunk;
  if (c < full) return (c * kXcds + x) * chunk + o;
  return full * kXcds * chunk + x * (per - full * chunk) + o;
}

// error bits raised by kernels (mapped to mfree_status on the host)
enum : uint32_t {
  EB_CELLS = 1u,      // num_cell >= n, or a hash outside the grid
  EB_NBH = 2u,        // >= max_nbh neighbours under the reference predicate
  EB_NONFINITE = 4u,  // NaN / Inf in a derivative
  EB_NONEIGH = 8u,    // particle without neighbour
  EB_SKIN = 16u,      // a particle moved more than half the neighbour-list skin within a step,
                      // or the skin exceeds kSkinMax h_max
};

// Raises error bits: in the context word err[0] and, when berr is set, in
// the word of the body (berr[body]; per-body words of a context that
// isolates its bodies)
__device__ inline void raise_err(uint32_t* err, uint32_t* berr, uint32_t body, uint32_t bits) {
  atomicOr(err, bits);
  if (berr) atomicOr(berr + body, bits);
}

// Per-body constants, device resident.  Derived constants are computed on
// the host with the reference's formulas (src/simulation_data.cpp:32-47).
struct SceneConst {
  double rho0, K, G2, c0sq;        // G2 = 2*G, c0sq = c0*c0 with c0 = sqrt(K/rho0)
  double E, nu;                    // the stress models of the co-rotated and potential TL lists
  double wdeltap, corr_exp, mghn_eps;
  double alpha, beta, eta, xsph_eps;
  double dt, half_dt, dt6;         // dt6 = dt*1./6.
  double time;                     // time at the start of the current step
  double v_bc;                     // TL strip velocity
  uint32_t offset, n, cell_offset, pow4;
};

// Per-body cell geometry of the current stage (src/grid.h:83-115).
struct SceneGeom {
  double bbmin_x, bbmin_y, cell;
  uint32_t nx, ny, num_cell, err;
  uint32_t tpc, ntiles, pad0, pad1;  // pair tiles: per cell column, in the body
  double skin;                       // neighbour-list skin of the step (0: per-stage cells)
};

struct Chunk {
  uint32_t scene, begin, end, pad;
};

// Sorted per-stage pair inputs, FP64 mode: one 128-byte record per particle.
// Godunov bodies (src/godunov.h) store p in rho21, S_yy without the pressure
// in syy (the reference's Godunov stress tensor, src/godunov.h:72-73) and no
// artificial stress.
struct alignas(16) PairRec64 {
  double vx, vy, rho, rho21;  // rho21 = 1/(rho*rho)
  double qw, m, h, c;         // qw = m/rho, c = sqrt(K/rho)
  double sxx, sxy, syy;       // sigma = S - p I
  double Rxx, Rxy, Ryy;       // artificial stress
  double pad0, pad1;
};

// Mixed mode: 64-byte record of FP32 values; positions are stored relative
// to the particle's own cell corner (exact FP64 subtraction, then rounded).
// The stress enters the pair terms only as sigma / rho^2, which is stored
// (computed in FP64).
struct alignas(16) PairRec32 {
  float vx, vy, rho, pad2;
  float qw, m, h, c;
  float sxx_r, sxy_r, syy_r, pad0;  // sigma / rho^2
  float Rxx, Rxy, Ryy, pad1;
};

}  // namespace mfree
