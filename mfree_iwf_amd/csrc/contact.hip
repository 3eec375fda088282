// Penalty contact between bodies on gfx950: the outline of a body
// (src/boundary.h:26-91, updated by body::update_boundary, src/body.h:152-165),
// the contact force of contact_apply_body_to_boundary (src/contact.h:36-80)
// and the convex hull the drivers take as outline (src/convex_hull.h:62-95).
//
// The reference finds each slave particle's closest master outline point
// with a k-d tree (src/kd_tree.cpp:20-116).  Outlines are short (the convex
// hull of a body: hundreds to a few thousand points), so here every slave
// particle scans the whole outline, staged through LDS in tiles shared by a
// workgroup: the result is the exact nearest point (of points at exactly the
// same squared distance, the first in outline order; the k-d tree returns
// the first it meets, which can differ only for such exact ties).  The
// contact kernel writes every slave particle's force (the reference zeroes
// them all, then sets those inside), by global idx, which the stepping
// kernels read as f_c / m in the momentum equations.
#include <float.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "ctx.h"

namespace mfree {

namespace {

constexpr uint32_t kBndTile = 1024;  // outline points per LDS tile (32 KiB)

// global idx -> storage index of the current storage
__global__ __launch_bounds__(kBlock) void k_inv_gid(uint32_t n, const uint32_t* __restrict__ gid,
                                                    uint32_t* __restrict__ inv) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n) inv[gid[s]] = s;
}

// body::update_boundary: outline point k takes its particle's position
__global__ __launch_bounds__(kBlock) void k_bnd_gather(uint32_t nb, const uint32_t* __restrict__ bgid,
                                                       const uint32_t* __restrict__ inv, const double* __restrict__ x,
                                                       const double* __restrict__ y, double4* __restrict__ pts) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nb) return;
  const uint32_t s = inv[bgid[k]];
  pts[k] = make_double4(x[s], y[s], 0., 0.);
}

// boundary::update_normals (src/boundary.h:47-73): the mean of the outward
// normals of the edges to the cyclic neighbours, each normalised as
// glm::normalize(v) = v * (1/sqrt(dot(v, v)))
__global__ __launch_bounds__(kBlock) void k_bnd_normals(uint32_t nb, const uint2* __restrict__ rng,
                                                        double4* __restrict__ pts) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nb) return;
  const uint2 r = rng[k];
  const uint32_t i = k - r.x;
  const uint32_t left = r.x + (i == 0 ? r.y - 1 : i - 1);
  const uint32_t right = r.x + (i + 1 >= r.y ? 0 : i + 1);
  const double4 p = pts[k], pl = pts[left], pr = pts[right];
  const double tlx = pl.x - p.x, tly = pl.y - p.y;
  const double trx = pr.x - p.x, try_ = pr.y - p.y;
  const double alx = -tly, aly = tlx;
  const double il = 1. / sqrt(alx * alx + aly * aly);
  const double nlx = alx * il, nly = aly * il;
  const double arx = try_, ary = -trx;
  const double ir = 1. / sqrt(arx * arx + ary * ary);
  const double nrx = arx * ir, nry = ary * ir;
  pts[k].z = 0.5 * (nlx + nrx);
  pts[k].w = 0.5 * (nly + nry);
}

// contact_apply_body_to_boundary for the slave particles in storage range
// [s0, s0 + ns) against master outline points pts[0, nb)
__global__ __launch_bounds__(kBlock) void k_contact(uint32_t s0, uint32_t ns, const double* __restrict__ x,
                                                    const double* __restrict__ y, const double* __restrict__ m,
                                                    const uint32_t* __restrict__ gid, const double4* __restrict__ pts,
                                                    uint32_t nb, double dt, double* __restrict__ fcx,
                                                    double* __restrict__ fcy) {
  __shared__ double4 tile[kBndTile];
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = t < ns;
  const uint32_t s = s0 + (live ? t : 0);
  const double qx = x[s], qy = y[s];
  double best = DBL_MAX;
  uint32_t bk = 0;
  for (uint32_t k0 = 0; k0 < nb; k0 += kBndTile) {
    const uint32_t cnt = min(kBndTile, nb - k0);
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < cnt; k += blockDim.x) tile[k] = pts[k0 + k];
    __syncthreads();
    for (uint32_t k = 0; k < cnt; k++) {
      const double dx = qx - tile[k].x;
      const double dy = qy - tile[k].y;
      const double dist2 = dx * dx + dy * dy;
      if (dist2 < best) {
        best = dist2;
        bk = k0 + k;
      }
    }
  }
  if (!live) return;
  const double4 cpt = pts[bk];
  const double dx = qx - cpt.x;
  const double dy = qy - cpt.y;
  const double pdotn = dx * cpt.z + dy * cpt.w;
  double fx = 0., fy = 0.;
  if (pdotn <= 0.) {
    const double dt2 = dt * dt;
    const double alpha = 0.1;
    const double gN_nx = cpt.z * pdotn;
    const double gN_ny = cpt.w * pdotn;
    const double ms = m[s];
    fx = -(ms * gN_nx / dt2 * alpha);
    fy = -(ms * gN_ny / dt2 * alpha);
  }
  const uint32_t g = gid[s];
  fcx[g] = fx;
  fcy[g] = fy;
}

inline uint32_t nblk(uint64_t n) { return (uint32_t)((n + kBlock - 1) / kBlock); }

// Uploads the outlines (concatenated in body order) when they changed.
mfree_status bnd_upload(mfree_ctx* c) {
  if (!c->bnd_dirty) return MFREE_OK;
  std::vector<uint32_t> gid;
  std::vector<uint2> rng;
  c->bnd_off.assign(c->S + 1, 0);
  for (uint32_t b = 0; b < c->S; b++) {
    c->bnd_off[b] = (uint32_t)gid.size();
    const uint32_t first = (uint32_t)gid.size(), cnt = (uint32_t)c->bnd[b].size();
    for (uint32_t i : c->bnd[b]) {
      gid.push_back(c->off[b] + i);
      rng.push_back(make_uint2(first, cnt));
    }
  }
  c->bnd_off[c->S] = (uint32_t)gid.size();
  const uint32_t nb = (uint32_t)gid.size();
  if (nb > c->bnd_cap) {
    const uint32_t cap = std::max(nb, 2 * c->bnd_cap);
    HIPCHK(c, dalloc(c, &c->bnd_gid, cap));
    HIPCHK(c, dalloc(c, &c->bnd_rng, cap));
    HIPCHK(c, dalloc(c, &c->bnd_pts, cap));
    c->bnd_cap = cap;
  }
  if (nb) {
    HIPCHK(c, hipMemcpy(c->bnd_gid, gid.data(), nb * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->bnd_rng, rng.data(), nb * sizeof(uint2), hipMemcpyHostToDevice));
  }
  c->bnd_dirty = false;
  return MFREE_OK;
}

}  // namespace

void drop_graphs(mfree_ctx* c) {
  for (auto& g : c->gexec)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

mfree_status contact_alloc(mfree_ctx* c) {
  const size_t N = c->N;
  HIPCHK(c, dalloc(c, &c->fcx, N));
  HIPCHK(c, dalloc(c, &c->fcy, N));
  HIPCHK(c, dalloc(c, &c->inv_gid, N));
  if (!c->hfcx.empty()) {
    HIPCHK(c, hipMemcpy(c->fcx, c->hfcx.data(), N * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->fcy, c->hfcy.data(), N * sizeof(double), hipMemcpyHostToDevice));
  } else {
    HIPCHK(c, hipMemset(c->fcx, 0, N * sizeof(double)));
    HIPCHK(c, hipMemset(c->fcy, 0, N * sizeof(double)));
  }
  c->bnd.resize(c->S);
  c->bnd_off.assign(c->S + 1, 0);
  bool any = !c->hfcx.empty();
  for (const auto& b : c->bnd) any = any || !b.empty();
  c->fc_active = any;
  return MFREE_OK;
}

mfree_status contact_activate(mfree_ctx* c) {
  if (c->fc_active) return MFREE_OK;
  c->fc_active = true;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  drop_graphs(c);
  return MFREE_OK;
}

}  // namespace mfree

using namespace mfree;

extern "C" {

mfree_status mfree_set_boundary(mfree_ctx* c, uint32_t body, const uint32_t* idx, uint32_t n) {
  if (!c || body >= c->prm.size() || (n && !idx)) return MFREE_ERR_ARG;
  if (c->opt.pipeline == MFREE_PIPELINE_REFERENCE)
    return fail(c, MFREE_ERR_UNSUPPORTED, "contact forces on the reference-structured pipeline");
  const uint32_t nbody = c->off[body + 1] - c->off[body];
  for (uint32_t k = 0; k < n; k++)
    if (idx[k] >= nbody) return fail(c, MFREE_ERR_ARG, "outline index outside the body");
  if (c->bnd.size() < c->prm.size()) c->bnd.resize(c->prm.size());
  c->bnd[body].assign(idx, idx + n);
  c->bnd_dirty = true;
  // before the first compute call the outline is kept for commit
  return (n && c->committed) ? contact_activate(c) : MFREE_OK;
}

mfree_status mfree_update_boundaries(mfree_ctx* c) {
  if (!c) return MFREE_ERR_ARG;
  mfree_status st = commit(c);
  if (st) return st;
  HIPCHK(c, hipSetDevice(c->opt.device));
  if ((st = bnd_upload(c))) return st;
  const uint32_t nb = c->bnd_off.empty() ? 0 : c->bnd_off[c->S];
  if (!nb) return MFREE_OK;
  const int b = c->cur;
  hipLaunchKernelGGL(k_inv_gid, dim3(nblk(c->N)), dim3(kBlock), 0, c->stream, c->N, c->gid[b], c->inv_gid);
  hipLaunchKernelGGL(k_bnd_gather, dim3(nblk(nb)), dim3(kBlock), 0, c->stream, nb, c->bnd_gid, c->inv_gid,
                     c->y[b].f[0], c->y[b].f[1], c->bnd_pts);
  hipLaunchKernelGGL(k_bnd_normals, dim3(nblk(nb)), dim3(kBlock), 0, c->stream, nb, c->bnd_rng, c->bnd_pts);
  HIPCHK(c, hipGetLastError());
  return MFREE_OK;
}

mfree_status mfree_get_boundary(mfree_ctx* c, uint32_t body, double* pts, uint32_t cap, uint32_t* n) {
  if (!c || body >= c->prm.size()) return MFREE_ERR_ARG;
  mfree_status st = commit(c);
  if (st) return st;
  if ((st = bnd_upload(c))) return st;
  const uint32_t o = c->bnd_off[body], cnt = c->bnd_off[body + 1] - o;
  if (n) *n = cnt;
  if (!pts) return MFREE_OK;  // size query
  if (cap < cnt) return fail(c, MFREE_ERR_CAPACITY, "outline buffer too small");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (cnt) HIPCHK(c, hipMemcpy(pts, c->bnd_pts + o, cnt * sizeof(double4), hipMemcpyDeviceToHost));
  return MFREE_OK;
}

mfree_status mfree_contact(mfree_ctx* master, uint32_t master_body, mfree_ctx* slave, uint32_t slave_body) {
  if (!master || !slave || master_body >= master->prm.size() || slave_body >= slave->prm.size())
    return MFREE_ERR_ARG;
  if (master->opt.device != slave->opt.device)
    return fail(slave, MFREE_ERR_UNSUPPORTED, "contact between bodies on different devices");
  mfree_status st = commit(master);
  if (st) return st;
  if ((st = commit(slave))) return st;
  // src/contact.h:38: nothing happens unless both bodies have an outline
  if (master->bnd[master_body].empty() || slave->bnd[slave_body].empty()) return MFREE_OK;
  if (master->bnd_dirty) return fail(master, MFREE_ERR_ARG, "outline changed: call mfree_update_boundaries first");
  HIPCHK(slave, hipSetDevice(slave->opt.device));
  if ((st = contact_activate(slave))) return st;
  if (master != slave) {
    // the master's outline update is enqueued on its own stream
    hipEvent_t e = nullptr;
    HIPCHK(slave, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIPCHK(slave, hipEventRecord(e, master->stream));
    HIPCHK(slave, hipStreamWaitEvent(slave->stream, e, 0));
    HIPCHK(slave, hipEventDestroy(e));
  }
  const uint32_t mo = master->bnd_off[master_body], nb = master->bnd_off[master_body + 1] - mo;
  const uint32_t s0 = slave->off[slave_body], ns = slave->off[slave_body + 1] - s0;
  if (!ns || !nb) return MFREE_OK;
  const int b = slave->cur;
  hipLaunchKernelGGL(k_contact, dim3(nblk(ns)), dim3(kBlock), 0, slave->stream, s0, ns, slave->y[b].f[0],
                     slave->y[b].f[1], slave->m[b], slave->gid[b], master->bnd_pts + mo, nb,
                     slave->prm[slave_body].dt, slave->fcx, slave->fcy);
  HIPCHK(slave, hipGetLastError());
  return MFREE_OK;
}

mfree_status mfree_contact_all(mfree_ctx* c) {
  if (!c) return MFREE_ERR_ARG;
  mfree_status st = commit(c);
  if (st) return st;
  // src/mfree_iwf.cpp:91-97: every ordered pair (master i, slave j), i != j,
  // in this order; a later master overwrites a slave's forces
  for (uint32_t i = 0; i < c->S; i++)
    for (uint32_t j = 0; j < c->S; j++)
      if (i != j && (st = mfree_contact(c, i, c, j))) return st;
  return MFREE_OK;
}

uint32_t mfree_convex_hull(const double* x, const double* y, uint32_t n, uint32_t* idx_out) {
  // Andrew's monotone chain (A. M. Andrew, Inf. Process. Lett. 9 (1979) 216)
  // with the reference's conventions (src/convex_hull.h:62-95): points
  // ordered by x, then y; a point is removed from the chain only for a
  // clockwise turn (collinear points stay); the lower chain from the
  // leftmost point, then the upper chain back, the start not repeated.
  if (!x || !y || !idx_out || n == 0) return 0;
  std::vector<uint32_t> ord(n);
  for (uint32_t i = 0; i < n; i++) ord[i] = i;
  std::sort(ord.begin(), ord.end(),
            [x, y](uint32_t a, uint32_t b) { return x[a] < x[b] || (x[a] == x[b] && y[a] < y[b]); });
  auto turn = [x, y](uint32_t o, uint32_t a, uint32_t b) {
    return (x[a] - x[o]) * (y[b] - y[o]) - (y[a] - y[o]) * (x[b] - x[o]);
  };
  std::vector<uint32_t> chain(2 * (size_t)n + 1);
  size_t k = 0;
  for (uint32_t i = 0; i < n; i++) {
    while (k >= 2 && turn(chain[k - 2], chain[k - 1], ord[i]) < 0) k--;
    chain[k++] = ord[i];
  }
  const size_t lower = k + 1;
  for (uint32_t i = n - 1; i-- > 0;) {
    while (k >= lower && turn(chain[k - 2], chain[k - 1], ord[i]) < 0) k--;
    chain[k++] = ord[i];
  }
  const uint32_t m = k > 0 ? (uint32_t)(k - 1) : 0;
  for (uint32_t i = 0; i < m; i++) idx_out[i] = chain[i];
  return m;
}

}  // extern "C"
