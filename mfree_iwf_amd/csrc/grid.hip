// Cell-list construction on gfx950: the reference's grid (src/grid.h) as a
// counting sort in HBM.
//
//   bbox      two-level min/max reduction per body (src/grid.h:83-115)
//   geometry  exactly the reference's nudge / cell side / ceil (src/grid.h:100-114)
//   hash      ix*ny + iy with (unsigned) truncation (src/grid.h:35-41),
//             offset by the body's cell range -> one key space for all bodies
//   count     atomic histogram of keys, rank of each particle in its cell
//   scan      exclusive prefix sum -> cell_start (src/grid.h:48-81: the
//             reference's "first index per cell, empty cells fixed from the
//             right" table is exactly this exclusive scan)
//   scatter   perm[cell_start[key] + rank] = particle
//   cell sort within each cell, particles ordered by idx, so the order no
//             longer depends on atomic timing (deterministic sums)
#include "kernels.h"

#include <float.h>

namespace mfree {

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

template <typename T>
__device__ inline T wave_reduce_min(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}
template <typename T>
__device__ inline T wave_reduce_max(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// One workgroup per chunk: partial {min x, min y, max x, max y, max h}, and
// max |v|^2 into vpart when velocities are given.
__global__ __launch_bounds__(kBlock) void k_bbox_partial(const Chunk* __restrict__ ch,
                                                         const double* __restrict__ x,
                                                         const double* __restrict__ y,
                                                         const double* __restrict__ h,
                                                         double* __restrict__ part,
                                                         const double* __restrict__ vx,
                                                         const double* __restrict__ vy,
                                                         double* __restrict__ vpart) {
  const Chunk c = ch[blockIdx.x];
  double mnx = DBL_MAX, mny = DBL_MAX, mxx = -DBL_MAX, mxy = -DBL_MAX, mxh = -DBL_MAX, mxv = 0.0;
  for (uint32_t i = c.begin + threadIdx.x; i < c.end; i += blockDim.x) {
    double xi = x[i], yi = y[i];
    mnx = fmin(xi, mnx);
    mny = fmin(yi, mny);
    mxx = fmax(xi, mxx);
    mxy = fmax(yi, mxy);
    mxh = fmax(h[i], mxh);
    if (vx) mxv = fmax(vx[i] * vx[i] + vy[i] * vy[i], mxv);
  }
  mnx = wave_reduce_min(mnx);
  mny = wave_reduce_min(mny);
  mxx = wave_reduce_max(mxx);
  mxy = wave_reduce_max(mxy);
  mxh = wave_reduce_max(mxh);
  mxv = wave_reduce_max(mxv);
  __shared__ double sh[6][kBlock / 64];
  int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) {
    sh[0][w] = mnx;
    sh[1][w] = mny;
    sh[2][w] = mxx;
    sh[3][w] = mxy;
    sh[4][w] = mxh;
    sh[5][w] = mxv;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < kBlock / 64; k++) {
      sh[0][0] = fmin(sh[0][0], sh[0][k]);
      sh[1][0] = fmin(sh[1][0], sh[1][k]);
      sh[2][0] = fmax(sh[2][0], sh[2][k]);
      sh[3][0] = fmax(sh[3][0], sh[3][k]);
      sh[4][0] = fmax(sh[4][0], sh[4][k]);
      sh[5][0] = fmax(sh[5][0], sh[5][k]);
    }
    double* o = part + 5 * (size_t)blockIdx.x;
    for (int k = 0; k < 5; k++) o[k] = sh[k][0];
    if (vx) vpart[blockIdx.x] = sh[5][0];
  }
}

// One workgroup per body: combine the chunk partials and derive the cell
// geometry with the reference's arithmetic (src/grid.h:100-114).
__global__ __launch_bounds__(kBlock) void k_bbox_final(const uint32_t* __restrict__ chunk_first,
                                                       const SceneConst* sc,
                                                       const double* __restrict__ part, double kernel_width,
                                                       double cell_min, uint32_t tile_cells,
                                                       SceneGeom* __restrict__ geom, uint32_t* __restrict__ err,
                                                       uint32_t* __restrict__ berr,
                                                       const double* __restrict__ vpart, double skin_min,
                                                       double skin_vfac, SceneConst* advance) {
  const uint32_t s = blockIdx.x;
  if (advance && threadIdx.x == 0) advance[s].time = advance[s].time + advance[s].dt;
  double mnx = DBL_MAX, mny = DBL_MAX, mxx = -DBL_MAX, mxy = -DBL_MAX, mxh = -DBL_MAX, mxv = 0.0;
  for (uint32_t c = chunk_first[s] + threadIdx.x; c < chunk_first[s + 1]; c += blockDim.x) {
    const double* p = part + 5 * (size_t)c;
    mnx = fmin(p[0], mnx);
    mny = fmin(p[1], mny);
    mxx = fmax(p[2], mxx);
    mxy = fmax(p[3], mxy);
    mxh = fmax(p[4], mxh);
    if (vpart) mxv = fmax(vpart[c], mxv);
  }
  mnx = wave_reduce_min(mnx);
  mny = wave_reduce_min(mny);
  mxx = wave_reduce_max(mxx);
  mxy = wave_reduce_max(mxy);
  mxh = wave_reduce_max(mxh);
  mxv = wave_reduce_max(mxv);
  __shared__ double sh[6][kBlock / 64];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) {
    sh[0][w] = mnx;
    sh[1][w] = mny;
    sh[2][w] = mxx;
    sh[3][w] = mxy;
    sh[4][w] = mxh;
    sh[5][w] = mxv;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int k = 1; k < kBlock / 64; k++) {
    mnx = fmin(sh[0][k], mnx);
    mny = fmin(sh[1][k], mny);
    mxx = fmax(sh[2][k], mxx);
    mxy = fmax(sh[3][k], mxy);
    mxh = fmax(sh[4][k], mxh);
    mxv = fmax(sh[5][k], mxv);
  }
  mnx = fmin(sh[0][0], mnx);
  mny = fmin(sh[1][0], mny);
  mxx = fmax(sh[2][0], mxx);
  mxy = fmax(sh[3][0], mxy);
  mxh = fmax(sh[4][0], mxh);
  mxv = fmax(sh[5][0], mxv);
  mnx -= 1e-6;
  mny -= 1e-6;
  mxx += 1e-6;
  mxy += 1e-6;
  // neighbour-list binning: the skin covers twice the largest displacement
  // of a step with skin_vfac/2 times the fastest particle's speed
  const double skin = vpart ? fmax(skin_min * mxh, skin_vfac * sc[s].dt * sqrt(mxv)) : 0.0;
  const double dx = vpart ? 2.0 * mxh + skin : fmax(mxh * kernel_width, cell_min);
  const double lx = mxx - mnx;
  const double ly = mxy - mny;
  const double fnx = ceil(lx / dx), fny = ceil(ly / dx);
  const bool skin_ok = !(skin > kSkinMax * mxh);
  SceneGeom g;
  g.bbmin_x = mnx;
  g.bbmin_y = mny;
  g.cell = dx;
  g.err = 0;
  // num_cell < n is asserted by the reference (src/grid.h:114); the same
  // bound keeps every key inside the body's cell range here.
  const double ncell = fnx * fny;
  // a body that already failed (its error word is set: within this call of
  // the steps, or frozen in a context that isolates its bodies) gets the
  // one-cell fallback too, so its pair passes scan nothing
  const uint32_t failed = berr ? berr[s] : 0u;
  if (!(fnx >= 1.0 && fny >= 1.0 && ncell < (double)sc[s].n) || !skin_ok || failed) {
    const uint32_t bit = failed ? failed : skin_ok ? EB_CELLS : EB_SKIN;
    g.nx = 1;
    g.ny = 1;
    g.num_cell = 1;
    g.err = bit;
    if (!failed) raise_err(err, berr, s, bit);
  } else {
    g.nx = (unsigned int)fnx;
    g.ny = (unsigned int)fny;
    g.num_cell = g.nx * g.ny;
  }
  g.tpc = tile_cells ? (g.ny + tile_cells - 1) / tile_cells : 0;
  g.ntiles = g.nx * g.tpc;
  g.pad0 = g.pad1 = 0;
  g.skin = skin;
  geom[s] = g;
}

// Per particle (storage order): key = cell_offset + ix*ny + iy, and its rank
// within the cell.  Storage is kept in cell order, so neighbouring lanes
// mostly share a key: each run of equal keys inside a wave takes one atomic
// for the whole run and hands out consecutive ranks.
__global__ __launch_bounds__(kBlock) void k_hash_count(uint32_t n, const uint32_t* __restrict__ scene,
                                                       const double* __restrict__ x, const double* __restrict__ y,
                                                       const SceneConst* __restrict__ sc,
                                                       const SceneGeom* __restrict__ geom,
                                                       uint32_t* __restrict__ key, uint32_t* __restrict__ rank,
                                                       uint32_t* __restrict__ count, uint32_t* __restrict__ err,
                                                       uint32_t* __restrict__ berr) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  uint32_t k = 0xFFFFFFFFu;
  if (i < n) {
    const uint32_t s = scene[i];
    const SceneGeom g = geom[s];
    const double fx = (x[i] - g.bbmin_x) / g.cell;
    const double fy = (y[i] - g.bbmin_y) / g.cell;
    uint32_t ix = 0, iy = 0;
    if ((fx >= 0.0) && (fy >= 0.0) && (fx < (double)g.nx) && (fy < (double)g.ny)) {
      ix = (unsigned int)fx;
      iy = (unsigned int)fy;
    } else if (!g.err) {
      raise_err(err, berr, s, EB_CELLS);
    }
    k = sc[s].cell_offset + ix * g.ny + iy;
  }
  const uint32_t prev = __shfl_up(k, 1, 64);
  const bool head = (lane == 0) || (prev != k);
  const uint64_t heads = __ballot(head);
  const uint64_t upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1);
  const int hl = 63 - __clzll(heads & upto);          // head lane of my run
  const uint64_t after = (lane == 63) ? 0ull : (heads & ~upto);
  const int next = after ? __ffsll((long long)after) - 1 : 64;  // next run's head
  uint32_t base = 0;
  if (head && k != 0xFFFFFFFFu) base = atomicAdd(&count[k], (uint32_t)(next - lane));
  base = __shfl(base, hl, 64);
  if (i < n) {
    key[i] = k;
    rank[i] = base + (uint32_t)(lane - hl);
  }
}

// ---- exclusive scan over u32 (cell counts -> cell starts) ---------------

__device__ inline uint32_t wave_incl_scan(uint32_t v) {
  int l = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t t = __shfl_up(v, o, 64);
    if (l >= o) v += t;
  }
  return v;
}

// Block-wide exclusive scan of one value per thread; returns the block total.
__device__ inline uint32_t block_excl_scan(uint32_t v, uint32_t* excl) {
  __shared__ uint32_t wsum[kBlock / 64];
  uint32_t inc = wave_incl_scan(v);
  int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 63) wsum[w] = inc;
  __syncthreads();
  uint32_t base = 0, tot = 0;
  for (int k = 0; k < kBlock / 64; k++) {
    if (k < w) base += wsum[k];
    tot += wsum[k];
  }
  __syncthreads();
  *excl = base + inc - v;
  return tot;
}

// Tile sums: each workgroup sums kScanTile elements.
__global__ __launch_bounds__(kBlock) void k_scan_tiles(const uint32_t* __restrict__ in, uint32_t n,
                                                       uint32_t* __restrict__ tsum) {
  uint32_t base = blockIdx.x * kScanTile;
  uint32_t v = 0;
  for (uint32_t k = 0; k < kScanTile / kBlock; k++) {
    uint32_t i = base + threadIdx.x * (kScanTile / kBlock) + k;
    if (i < n) v += in[i];
  }
  uint32_t ex;
  uint32_t tot = block_excl_scan(v, &ex);
  if (threadIdx.x == 0) tsum[blockIdx.x] = tot;
}

// out[i] = tile base + exclusive scan within the tile; out[n] = total.
// Each workgroup sums the tile sums before its own (at most a few thousand
// values).  clear: zero in[] after reading it (the binning's cell counts,
// which then start every binning at zero without a fill pass), or null.
__global__ __launch_bounds__(kBlock) void k_scan_apply(uint32_t* __restrict__ in, uint32_t n,
                                                       const uint32_t* __restrict__ tsum,
                                                       uint32_t* __restrict__ out, uint32_t total, int clear) {
  constexpr uint32_t per = kScanTile / kBlock;
  uint32_t pre = 0;
  for (uint32_t t = threadIdx.x; t < blockIdx.x; t += kBlock) pre += tsum[t];
  uint32_t ex;
  const uint32_t tile_base = block_excl_scan(pre, &ex);
  uint32_t base = blockIdx.x * kScanTile + threadIdx.x * per;
  uint32_t vals[per];
  uint32_t v = 0;
#pragma unroll
  for (uint32_t k = 0; k < per; k++) {
    uint32_t i = base + k;
    vals[k] = i < n ? in[i] : 0;
    v += vals[k];
  }
  if (clear) {
#pragma unroll
    for (uint32_t k = 0; k < per; k++)
      if (base + k < n) in[base + k] = 0;
  }
  block_excl_scan(v, &ex);
  uint32_t run = tile_base + ex;
#pragma unroll
  for (uint32_t k = 0; k < per; k++) {
    uint32_t i = base + k;
    if (i < n) out[i] = run;
    run += vals[k];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = total;
}

// One workgroup: first pair tile of every body (exclusive scan of the
// per-body tile counts); tile_off[nscene] = total.
__global__ __launch_bounds__(kBlock) void k_tile_offsets(uint32_t nscene, const SceneGeom* __restrict__ geom,
                                                         uint32_t* __restrict__ tile_off) {
  uint32_t carry = 0;
  for (uint32_t b = 0; b < nscene; b += kBlock) {
    const uint32_t s = b + threadIdx.x;
    const uint32_t v = s < nscene ? geom[s].ntiles : 0;
    uint32_t ex;
    const uint32_t tot = block_excl_scan(v, &ex);
    if (s < nscene) tile_off[s] = carry + ex;
    carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) tile_off[nscene] = carry;
}

__global__ __launch_bounds__(kBlock) void k_scatter(uint32_t n, const uint32_t* __restrict__ key,
                                                    const uint32_t* __restrict__ rank,
                                                    const uint32_t* __restrict__ cell_start,
                                                    uint32_t* __restrict__ perm) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  perm[cell_start[key[i]] + rank[i]] = i;
}

// The scatter of a deterministic binning with a rank sort: besides the
// particle, its global id and cell key (bit 31: its body's grid fell back to
// one cell) go to the particle's scatter slot, so that k_cell_rank reads a
// slot's data and its cell-mates' ids without going through the particle.
__global__ __launch_bounds__(kBlock) void k_scatter_slots(uint32_t n, const uint32_t* __restrict__ key,
                                                          const uint32_t* __restrict__ rank,
                                                          const uint32_t* __restrict__ cell_start,
                                                          const uint32_t* __restrict__ gid,
                                                          const uint32_t* __restrict__ scene,
                                                          const SceneGeom* __restrict__ geom,
                                                          uint32_t* __restrict__ perm, uint32_t* __restrict__ gid_at,
                                                          uint32_t* __restrict__ key_at) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t k = key[i];
  const uint32_t t = cell_start[k] + rank[i];
  perm[t] = i;
  gid_at[t] = gid[i];
  key_at[t] = k | (geom[scene[i]].err ? 0x80000000u : 0u);
}

// One thread per cell: insertion sort of the cell's entries by global id,
// so that within a cell particles appear in idx order.  Entries arrive
// nearly sorted (storage is in the previous step's cell order), so a cell
// costs about its size; a body whose grid fell back to one cell (its step
// fails) is left unsorted, since sorting that cell is quadratic in the body.
__global__ __launch_bounds__(kBlock) void k_cell_sort(uint32_t ncell, const uint32_t* __restrict__ cell_start,
                                                      const uint32_t* __restrict__ gid,
                                                      const uint32_t* __restrict__ scene,
                                                      const SceneGeom* __restrict__ geom,
                                                      uint32_t* __restrict__ perm) {
  uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  uint32_t a = cell_start[c], b = cell_start[c + 1];
  if (b - a < 2) return;
  if (geom[scene[perm[a]]].err) return;
  for (uint32_t i = a + 1; i < b; i++) {
    uint32_t p = perm[i];
    uint32_t g = gid[p];
    uint32_t j = i;
    while (j > a) {
      uint32_t q = perm[j - 1];
      if (gid[q] <= g) break;
      perm[j] = q;
      j--;
    }
    perm[j] = p;
  }
}

// One thread per sorted slot after the scatter into perm_in: the entry's
// rank among its cell's entries by global id (ties, which do not occur for
// distinct particles, by slot) places it in perm, so that within a cell the
// particles appear in idx order -- the order k_cell_sort produces.  Every
// load is independent of the comparisons, where an insertion sort chains
// dependent loads.  A body whose grid fell back to one cell (its step fails)
// keeps the scatter's order: ranking that cell is quadratic in the body.
__global__ __launch_bounds__(kBlock) void k_cell_rank(uint32_t n, const uint32_t* __restrict__ cell_start,
                                                      const uint32_t* __restrict__ gid_at,
                                                      const uint32_t* __restrict__ key_at,
                                                      const uint32_t* __restrict__ perm_in,
                                                      uint32_t* __restrict__ perm) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t p = perm_in[i], kf = key_at[i], g = gid_at[i];
  if (kf & 0x80000000u) {
    perm[i] = p;
    return;
  }
  const uint32_t a = cell_start[kf], b = cell_start[kf + 1];
  uint32_t r = 0;
  for (uint32_t j = a; j < b; j++) {
    const uint32_t gj = gid_at[j];
    r += (gj < g || (gj == g && j < i)) ? 1u : 0u;
  }
  perm[a + r] = p;
}

// Plain streaming copy: the HBM calibration kernel (reads and writes n*8 bytes).
__global__ __launch_bounds__(kBlock) void k_calib_copy(const double2* __restrict__ a, double2* __restrict__ b,
                                                       uint64_t n2) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n2; i += (uint64_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}

// out[n] = out[n-1] + in[n-1]: total of an exclusive scan (n >= 1).
__global__ void k_scan_total(const uint32_t* __restrict__ in, uint32_t n, uint32_t* __restrict__ out) {
  if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = out[n - 1] + in[n - 1];
}

// ---- host launchers -------------------------------------------------------

static inline uint32_t nblk(uint64_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }

void launch_fill_u32(hipStream_t st, uint32_t* p, uint32_t v, uint32_t n) {
  if (n) hipLaunchKernelGGL(k_fill_u32, dim3(nblk(n, kBlock)), dim3(kBlock), 0, st, p, v, n);
}

void launch_calib_copy(hipStream_t st, const double* a, double* b, uint64_t n) {
  hipLaunchKernelGGL(k_calib_copy, dim3(256 * 16), dim3(kBlock), 0, st, (const double2*)a, (double2*)b, n / 2);
}

void launch_excl_scan(hipStream_t st, const uint32_t* in, uint32_t n, uint32_t* tsum, uint32_t* out) {
  if (n == 0) {
    launch_fill_u32(st, out, 0u, 1);
    return;
  }
  const uint32_t nt = nblk(n, kScanTile);
  hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(kBlock), 0, st, in, n, tsum);
  hipLaunchKernelGGL(k_scan_apply, dim3(nt), dim3(kBlock), 0, st, const_cast<uint32_t*>(in), n, tsum, out, 0u, 0);
  hipLaunchKernelGGL(k_scan_total, dim3(1), dim3(64), 0, st, in, n, out);
}

void launch_binning(hipStream_t st, const BinArgs& a) {
  const bool skin = a.vx && a.vy && a.vpart;
  hipLaunchKernelGGL(k_bbox_partial, dim3(a.nchunks), dim3(kBlock), 0, st, a.chunks, a.x, a.y, a.h, a.part,
                     skin ? a.vx : nullptr, skin ? a.vy : nullptr, skin ? a.vpart : nullptr);
  hipLaunchKernelGGL(k_bbox_final, dim3(a.nscene), dim3(kBlock), 0, st, a.chunk_first, a.sc, a.part,
                     a.kernel_width, a.cell_min, a.tile_cells, a.geom, a.err, a.berr, skin ? a.vpart : nullptr,
                     a.skin_min, a.skin_vfac, a.advance);
  if (a.tile_cells && a.tile_off)
    hipLaunchKernelGGL(k_tile_offsets, dim3(1), dim3(kBlock), 0, st, a.nscene, a.geom, a.tile_off);
  // a.count is zero here: zeroed at allocation and cleared again by the scan
  hipLaunchKernelGGL(k_hash_count, dim3(nblk(a.n, kBlock)), dim3(kBlock), 0, st, a.n, a.scene, a.x, a.y, a.sc,
                     a.geom, a.key, a.rank, a.count, a.err, a.berr);
  uint32_t nt = nblk(a.ncell_cap, kScanTile);
  hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(kBlock), 0, st, a.count, a.ncell_cap, a.tsum);
  hipLaunchKernelGGL(k_scan_apply, dim3(nt), dim3(kBlock), 0, st, a.count, a.ncell_cap, a.tsum, a.cell_start,
                     a.n, 1);
  const bool ranked = a.deterministic && a.perm_tmp;
  if (ranked) {
    uint32_t* gid_at = a.perm_tmp + a.n;
    uint32_t* key_at = a.perm_tmp + 2 * (size_t)a.n;
    hipLaunchKernelGGL(k_scatter_slots, dim3(nblk(a.n, kBlock)), dim3(kBlock), 0, st, a.n, a.key, a.rank,
                       a.cell_start, a.gid, a.scene, a.geom, a.perm_tmp, gid_at, key_at);
    hipLaunchKernelGGL(k_cell_rank, dim3(nblk(a.n, kBlock)), dim3(kBlock), 0, st, a.n, a.cell_start, gid_at,
                       key_at, a.perm_tmp, a.perm);
  } else {
    hipLaunchKernelGGL(k_scatter, dim3(nblk(a.n, kBlock)), dim3(kBlock), 0, st, a.n, a.key, a.rank,
                       a.cell_start, a.perm);
  }
  if (!ranked && a.deterministic)
    hipLaunchKernelGGL(k_cell_sort, dim3(nblk(a.ncell_cap, kBlock)), dim3(kBlock), 0, st, a.ncell_cap,
                       a.cell_start, a.gid, a.scene, a.geom, a.perm);
}

}  // namespace mfree
